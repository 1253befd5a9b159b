// ORACLE (test infrastructure): C ABI over the reference's shipped, pre-generated
// old-API Hydro class (examples-old/Hydro-exampled/dist/Hydro.hpp:15-124).
#include "Hydro.hpp"
extern "C" {
void* ref_create() { return new Hydro(); }
void ref_destroy(void* h) { delete static_cast<Hydro*>(h); }
void ref_run(void* h, int k) {
  Hydro* m = static_cast<Hydro*>(h);
  if (k == 0) m->init(); else m->proceed();
}
void* ref_static_ptr(void* h, int s) {
  Hydro* m = static_cast<Hydro*>(h);
  switch (s) {
    case 0: return &m->static_0_generation;
    case 1: return &m->static_1_time;
    case 2: return &m->static_2_cfl;
    case 3: return &m->static_3_dR0;
    case 4: return &m->static_4_dR1;
    case 5: return &m->static_5_extent0;
    case 6: return &m->static_6_extent1;
    case 7: return m->static_7_density.data();
    case 8: return m->static_8_velocity0.data();
    case 9: return m->static_9_velocity1.data();
    case 10: return m->static_10_pressure.data();
  }
  return nullptr;
}
int ref_memory_size() { return 1060900; }
}
