// ORACLE (test infrastructure): C ABI over the reference's shipped, pre-generated
// old-API Life class (examples-old/Life-exampled/dist/Life.hpp:15-66).
#include "Life.hpp"
extern "C" {
void* ref_create() { return new Life(); }
void ref_destroy(void* h) { delete static_cast<Life*>(h); }
void ref_run(void* h, int k) {
  Life* m = static_cast<Life*>(h);
  if (k == 0) m->init(); else m->proceed();
}
void* ref_static_ptr(void* h, int s) {
  Life* m = static_cast<Life*>(h);
  switch (s) {
    case 0: return &m->static_0_population;
    case 1: return &m->static_1_generation;
    case 2: return m->static_2_cell.data();
  }
  return nullptr;
}
int ref_memory_size() { Life m; return m.memorySize(); }
}
