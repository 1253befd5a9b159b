"""Emitter of the generated C++ host class for a HIP machine.

The reference emits a class per OM with a constructor, one ``void`` method per
kernel, ``T& name()`` / ``T& name(i0, i1, ...)`` accessors and ``om_size*``
queries (``Generator/PlanTrans.hs:50-215``).  Storage here is device-resident,
so accessors cannot hand out host references: scalars are read with
``name()`` and written with ``set_name(v)``, arrays are read into and written
from ``std::vector`` (x-fastest), and ``name_devptr()`` exposes the device
buffer.  The class loads the machine library with ``dlopen`` and drives it
through ``include/paraiso_hip.h``.
"""
from __future__ import annotations

from ...om.dyn import Realm


def emit_host_class(om, setup, margins=None) -> str:
    name = om.name
    L = []
    w = L.append
    w("// Generated host class for Orthotope Machine '%s' (HIP, gfx950)." % name)
    w("#pragma once")
    w("#include <dlfcn.h>")
    w("#include <stdexcept>")
    w("#include <string>")
    w("#include <vector>")
    w('#include "paraiso_hip.h"')
    w("")
    w("class %s {" % name)
    w("  pr_machine* m_ = nullptr;")
    w("  void* lib_ = nullptr;")
    w("  static void check(int rc, pr_machine* m, const char* what) {")
    w("    if (rc != 0) throw std::runtime_error(std::string(what) + \": \" + (m ? pr_last_error(m) : \"\"));")
    w("  }")
    w(" public:")
    sizes = ", ".join(str(x) for x in setup.local_size)
    w("  explicit %s(const char* machine_library = \"lib%s.so\") {" % (name, name))
    w("    lib_ = dlopen(machine_library, RTLD_NOW | RTLD_LOCAL);")
    w("    if (!lib_) throw std::runtime_error(dlerror());")
    w("    auto describe = (const pr_machine_desc* (*)(void))dlsym(lib_, \"pm_describe\");")
    w("    if (!describe) throw std::runtime_error(\"pm_describe missing\");")
    w("    const int n[3] = {%s};" % ", ".join(str(x) for x in list(setup.local_size) + [1] * (3 - len(setup.local_size))))
    w("    check(pr_create(describe(), n, &m_), nullptr, \"pr_create\");")
    w("  }")
    w("  ~%s() { if (m_) pr_destroy(m_); if (lib_) dlclose(lib_); }" % name)
    w("  %s(const %s&) = delete;" % (name, name))
    w("  %s& operator=(const %s&) = delete;" % (name, name))
    w("  pr_machine* handle() { return m_; }")
    lm, um = margins if margins else ((0,) * setup.ndim, (0,) * setup.ndim)
    total = mem_total = 1
    for ax, x in enumerate(setup.local_size):
        total *= x
        mem_total *= x + lm[ax] + um[ax]
    w("  int om_size() const { return %d; }" % total)
    w("  int om_memory_size() const { return %d; }" % mem_total)
    for ax, x in enumerate(setup.local_size):
        w("  int om_size_%d() const { return %d; }" % (ax, x))
        w("  int om_memory_size_%d() const { return %d; }" % (ax, x + lm[ax] + um[ax]))
        w("  int om_lower_margin_%d() const { return %d; }" % (ax, lm[ax]))
        w("  int om_upper_margin_%d() const { return %d; }" % (ax, um[ax]))
    for ki, k in enumerate(om.kernels):
        w("  void %s() { check(pr_run(m_, %d, 1), m_, \"%s\"); }" % (k.name, ki, k.name))
    w("  void save(const char* path) { check(pr_state_save(m_, path), m_, \"save\"); }")
    w("  void load(const char* path) { check(pr_state_load(m_, path), m_, \"load\"); }")
    for si, s in enumerate(om.setup.static_values):
        t = s.dyn.ctype.c_name
        if s.dyn.realm is Realm.Scalar:
            w("  %s %s() { %s v; check(pr_static_read(m_, %d, &v, sizeof v, 0), m_, \"read\"); return v; }"
              % (t, s.name, t, si))
            w("  void set_%s(%s v) { check(pr_static_write(m_, %d, &v, sizeof v, 0), m_, \"write\"); }"
              % (s.name, t, si))
        else:
            w("  std::vector<%s> %s() { std::vector<%s> v(om_memory_size()); check(pr_static_read(m_, %d, v.data(), "
              "v.size() * sizeof(%s), 0), m_, \"read\"); return v; }" % (t, s.name, t, si, t))
            w("  void set_%s(const std::vector<%s>& v) { check(pr_static_write(m_, %d, v.data(), "
              "v.size() * sizeof(%s), 0), m_, \"write\"); }" % (s.name, t, si, t))
            w("  void* %s_devptr() { void* p; check(pr_static_devptr(m_, %d, &p), m_, \"devptr\"); return p; }"
              % (s.name, si))
    w("};")
    return "\n".join(L) + "\n"
