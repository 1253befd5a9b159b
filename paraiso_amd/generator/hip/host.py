"""Emitter of the generated C++ host class for a HIP machine.

The reference emits a class per OM with a constructor, one ``void`` method per
kernel, ``T& name()`` / ``T& name(i0, i1, ...)`` accessors and ``om_size*``
queries (``Generator/PlanTrans.hs:50-215``); its CUDA flavour mirrors each
storage between host and device with a lazy-sync flag (``PlanTrans.hs:728``).

Here storage is device-resident and every static has a host mirror with the
same lazy synchronisation:

* ``T& name()`` and, for arrays, ``T& name(i0, i1, ...)`` (local coordinates,
  the reference's margin offsets, ``PlanTrans.hs:117-157``) bring the mirror
  up to date if the device holds newer values and mark the mirror as the newer
  copy, because the caller may write through the reference;
* a kernel call first writes back every mirror marked newer, then marks the
  statics the kernel writes as newer on the device;
* ``set_name(...)`` writes the device directly, ``name_devptr()`` hands out the
  device buffer (after writing back the mirror), and ``save``/``load`` write
  and read whole-machine state files.

The class loads the machine library with ``dlopen`` and drives it through
``include/paraiso_hip.h``; it is host code only.
"""
from __future__ import annotations

from ...om.dyn import CType, Realm
from ...om.graph import Store


def _elem(s) -> str:
    """Host element type: Bool arrays are bytes (std::vector<bool> has no data())."""
    return "unsigned char" if s.dyn.ctype is CType.Bool else s.dyn.ctype.c_name


def emit_host_class(om, setup, margins=None) -> str:
    name = om.name
    statics = om.setup.static_values
    lm, um = margins if margins else ((0,) * setup.ndim, (0,) * setup.ndim)
    mem = [x + lm[ax] + um[ax] for ax, x in enumerate(setup.local_size)]
    total = mem_total = 1
    for ax, x in enumerate(setup.local_size):
        total *= x
        mem_total *= mem[ax]
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
    w("  // mirror state per static: 0 in sync, 1 host newer, 2 device newer")
    w("  int state_[%d] = {0};" % max(1, len(statics)))
    for si, s in enumerate(statics):
        t = _elem(s) if s.dyn.realm is Realm.Array else s.dyn.ctype.c_name
        if s.dyn.realm is Realm.Array:
            w("  std::vector<%s> h%d_ = std::vector<%s>(%d);" % (t, si, t, mem_total))
        else:
            w("  %s h%d_ = %s(0);" % (t, si, t))
    w("  static void check(int rc, pr_machine* m, const char* what) {")
    w("    if (rc != 0) throw std::runtime_error(std::string(what) + \": \" + (m ? pr_last_error(m) : \"\"));")
    w("  }")
    w("  void pull(int s, void* h, size_t bytes) {")
    w("    if (state_[s] == 2) { check(pr_static_read(m_, s, h, bytes, 0), m_, \"read\"); state_[s] = 0; }")
    w("  }")
    w("  void push(int s, const void* h, size_t bytes) {")
    w("    if (state_[s] == 1) { check(pr_static_write(m_, s, h, bytes, 0), m_, \"write\"); state_[s] = 0; }")
    w("  }")
    w("  void push_all() {")
    for si, s in enumerate(statics):
        w("    push(%d, %s, %s);" % (si, _ptr(si, s), _bytes(si, s)))
    w("  }")
    w(" public:")
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
    w("  pr_machine* handle() { push_all(); return m_; }")
    w("  int om_size() const { return %d; }" % total)
    w("  int om_memory_size() const { return %d; }" % mem_total)
    for ax, x in enumerate(setup.local_size):
        w("  int om_size_%d() const { return %d; }" % (ax, x))
        w("  int om_memory_size_%d() const { return %d; }" % (ax, mem[ax]))
        w("  int om_lower_margin_%d() const { return %d; }" % (ax, lm[ax]))
        w("  int om_upper_margin_%d() const { return %d; }" % (ax, um[ax]))
    for ki, k in enumerate(om.kernels):
        written = sorted(set(nd.inst.static for nd in k.graph.nodes if isinstance(nd.inst, Store)))
        marks = " ".join("state_[%d] = 2;" % s for s in written)
        w("  void %s() { push_all(); check(pr_run(m_, %d, 1), m_, \"%s\"); %s }"
          % (k.name, ki, k.name, marks))
    w("  void save(const char* path) { push_all(); check(pr_state_save(m_, path), m_, \"save\"); }")
    w("  void load(const char* path) {")
    w("    check(pr_state_load(m_, path), m_, \"load\");")
    w("    for (int s = 0; s < %d; ++s) state_[s] = 2;" % len(statics))
    w("  }")
    for si, s in enumerate(statics):
        t = _elem(s) if s.dyn.realm is Realm.Array else s.dyn.ctype.c_name
        if s.dyn.realm is Realm.Scalar:
            w("  %s& %s() { pull(%d, &h%d_, sizeof h%d_); state_[%d] = 1; return h%d_; }"
              % (t, s.name, si, si, si, si, si))
            w("  void set_%s(%s v) { h%d_ = v; state_[%d] = 1; push(%d, &h%d_, sizeof h%d_); }"
              % (s.name, t, si, si, si, si, si))
            continue
        nbytes = "h%d_.size() * sizeof(%s)" % (si, t)
        w("  std::vector<%s>& %s() { pull(%d, h%d_.data(), %s); state_[%d] = 1; return h%d_; }"
          % (t, s.name, si, si, nbytes, si, si))
        args = ", ".join("int i%d" % ax for ax in range(setup.ndim))
        idx = "(%d + i%d)" % (lm[setup.ndim - 1], setup.ndim - 1)
        for ax in range(setup.ndim - 2, -1, -1):
            idx = "(%d + i%d) + %d * %s" % (lm[ax], ax, mem[ax], idx)
        w("  %s& %s(%s) { pull(%d, h%d_.data(), %s); state_[%d] = 1; return h%d_[%s]; }"
          % (t, s.name, args, si, si, nbytes, si, si, idx))
        w("  void set_%s(const std::vector<%s>& v) {" % (s.name, t))
        w("    if (v.size() != h%d_.size()) throw std::runtime_error(\"set_%s: size\");" % (si, s.name))
        w("    h%d_ = v; state_[%d] = 1; push(%d, h%d_.data(), %s);" % (si, si, si, si, nbytes))
        w("  }")
        w("  void* %s_devptr() { push(%d, h%d_.data(), %s); void* p; "
          "check(pr_static_devptr(m_, %d, &p), m_, \"devptr\"); state_[%d] = 2; return p; }"
          % (s.name, si, si, nbytes, si, si))
    w("};")
    return "\n".join(L) + "\n"


def _ptr(si, s) -> str:
    return ("h%d_.data()" % si) if s.dyn.realm is Realm.Array else ("&h%d_" % si)


def _bytes(si, s) -> str:
    if s.dyn.realm is Realm.Array:
        return "h%d_.size() * sizeof(%s)" % (si, _elem(s))
    return "sizeof h%d_" % si
