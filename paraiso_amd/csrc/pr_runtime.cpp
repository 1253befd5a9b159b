// Orthotope Machine runtime for MI355X: the hand-written side of the C ABI in
// include/paraiso_hip.h.  It owns device storage and executes the launch
// sequence a generated machine describes.
//
// Storage model (replacing the reference's std::vector/thrust_vector members
// and deep-copy stores, Generator/PlanTrans.hs:75-86,243-258):
//   * every Array static has one device buffer, plus a second one when some
//     kernel stores it; a store writes the second buffer and the pair is
//     swapped after the kernel, so every launch of the kernel still reads the
//     old value, as the reference's copy-after-all-subkernels does;
//   * Scalar statics, tail-computed scalar slots and arrival counters share
//     one small device block (8 bytes per entry), so a step never returns
//     to the host;
//   * reduction partials live in one buffer sized from the tile counts.
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <array>
#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>
#include <string>
#include <vector>

#include "paraiso_hip.h"

// A captured graph is bound to the buffers it was captured on: the ping-pong parity of
// every static (of every grid, for a batch) is part of its key.
struct GraphKey {
  int kernel, count;
  std::vector<char> parity;
  bool operator<(const GraphKey& o) const {
    if (kernel != o.kernel) return kernel < o.kernel;
    if (count != o.count) return count < o.count;
    return parity < o.parity;
  }
};

struct pr_machine {
  const pr_machine_desc* d = nullptr;
  int n[3] = {1, 1, 1};              // array extents (local size plus margins)
  int local[3] = {1, 1, 1};          // local size (the reference's localSize)
  long long cells = 1;               // cells of one static array
  std::vector<void*> buf[2];        // per static
  std::vector<int> cur;             // parity per static
  std::vector<char> stored_any;     // static stored by some kernel
  char* scal = nullptr;
  size_t scal_bytes = 0;
  char* part = nullptr;
  std::vector<std::vector<size_t>> part_off;    // [kernel][launch]
  std::vector<std::vector<unsigned>> ntiles;    // [kernel][launch]
  std::vector<std::vector<std::array<int, 3>>> tiles;
  std::vector<std::vector<size_t>> counter_idx; // [kernel][launch]
  hipStream_t stream = nullptr;
  bool own_stream = false;
  std::map<GraphKey, hipGraphExec_t> graphs;
  std::vector<std::vector<void*>> snaps;        // snapshot buffers
  std::vector<char*> snap_scal;
  std::vector<int> snap_sel;                    // statics a snapshot holds
  std::string err;
  hipEvent_t* timing = nullptr;    // set while pr_run_timed records launch events
  int timing_stride = 0, timing_call = 0;
};

namespace {

size_t ctype_size(int ct) {
  switch (ct) {
    case PR_BOOL: return 1;
    case PR_INT: return 4;
    case PR_FLOAT: return 4;
    case PR_DOUBLE: return 8;
  }
  return 8;
}

int fail(pr_machine* m, int code, const std::string& msg) {
  if (m) m->err = msg;
  return code;
}

int hip_fail(pr_machine* m, hipError_t e, const char* what) {
  return fail(m, PR_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define PR_HIP(m, call)                                    \
  do {                                                     \
    hipError_t e_ = (call);                                \
    if (e_ != hipSuccess) return hip_fail((m), e_, #call); \
  } while (0)

std::vector<char> parity_key(const pr_machine* m) {
  std::vector<char> p(m->cur.size());
  for (size_t s = 0; s < m->cur.size(); ++s) p[s] = (char)(m->cur[s] & 1);
  return p;
}

// One call of kernel k.  A static with a non-null entry in cur_ov is read from that buffer
// instead of the machine's own, and a store of it goes to next_ov[s]; such statics keep
// their ping-pong parity (the adjoint driver's tape lives outside the machine's buffers).
int launch_kernel_ov(pr_machine* m, int k, const void* const* cur_ov, void* const* next_ov) {
  const pr_kernel_desc& kd = m->d->kernels[k];
  const int ns = m->d->n_static;
  std::vector<const void*> cur(ns, nullptr);
  std::vector<void*> next(ns, nullptr);
  for (int s = 0; s < ns; ++s) {
    if (m->d->statics[s].realm != PR_ARRAY) continue;
    if (cur_ov && cur_ov[s]) {
      cur[s] = cur_ov[s];
      next[s] = next_ov ? next_ov[s] : nullptr;
      continue;
    }
    cur[s] = m->buf[m->cur[s]][s];
    if (m->stored_any[s]) next[s] = m->buf[1 - m->cur[s]][s];
  }
  for (int l = 0; l < kd.n_launch; ++l) {
    pr_launch_ctx ctx;
    ctx.cur = cur.data();
    ctx.next = next.data();
    ctx.scal = m->scal;
    ctx.part = m->part + m->part_off[k][l];
    ctx.counter = reinterpret_cast<unsigned*>(m->scal + 8 * m->counter_idx[k][l]);
    for (int a = 0; a < 3; ++a) {
      ctx.n[a] = m->n[a];
      ctx.tiles[a] = m->tiles[k][l][a];
    }
    ctx.ntiles = m->ntiles[k][l];
    ctx.stream = m->stream;
    hipEvent_t* ev = m->timing ? m->timing + 2 * ((size_t)m->timing_call * m->timing_stride + l)
                               : nullptr;
    if (ev) (void)hipEventRecord(ev[0], m->stream);
    int rc = kd.launches[l].fn(&ctx);
    if (ev) (void)hipEventRecord(ev[1], m->stream);
    if (rc != 0) {
      hipError_t e = hipGetLastError();
      return fail(m, PR_E_HIP, std::string("launch ") + kd.name + "[" + std::to_string(l) +
                                   "]: " + hipGetErrorString(e));
    }
  }
  for (int i = 0; i < kd.n_stored; ++i) {
    const int s = kd.stored[i];
    if (!(cur_ov && cur_ov[s])) m->cur[s] ^= 1;
  }
  return PR_OK;
}

int launch_kernel_once(pr_machine* m, int k) { return launch_kernel_ov(m, k, nullptr, nullptr); }

}  // namespace

extern "C" {

int pr_create(const pr_machine_desc* d, const int* local_size, pr_machine** out) {
  if (!d || !local_size || !out) return PR_E_INVALID;
  if (d->abi != PR_ABI_VERSION) return PR_E_ABI;
  *out = nullptr;
  pr_machine* m = new pr_machine();
  m->d = d;
  for (int a = 0; a < d->ndim; ++a) {
    if (local_size[a] < 1 || local_size[a] < d->halo) {
      int rc = fail(m, PR_E_INVALID, "grid extent " + std::to_string(local_size[a]) +
                                          " on axis " + std::to_string(a) +
                                          " is below the machine's halo " + std::to_string(d->halo));
      delete m;
      return rc;
    }
    m->local[a] = local_size[a];
    m->n[a] = local_size[a] + d->lower_margin[a] + d->upper_margin[a];
    m->cells *= m->n[a];
  }
  if (m->cells >= (1LL << 31)) {   // generated kernels address cells with 32-bit offsets
    int rc = fail(m, PR_E_INVALID, "grid of " + std::to_string(m->cells) +
                                       " cells exceeds 2^31 - 1 cells");
    delete m;
    return rc;
  }
  const int ns = d->n_static;
  m->cur.assign(ns, 0);
  m->stored_any.assign(ns, 0);
  m->buf[0].assign(ns, nullptr);
  m->buf[1].assign(ns, nullptr);
  for (int k = 0; k < d->n_kernel; ++k)
    for (int i = 0; i < d->kernels[k].n_stored; ++i) m->stored_any[d->kernels[k].stored[i]] = 1;
  // scalar block: statics, slots, counters (8 bytes each)
  size_t n_counter = 0;
  m->counter_idx.resize(d->n_kernel);
  for (int k = 0; k < d->n_kernel; ++k)
    for (int l = 0; l < d->kernels[k].n_launch; ++l)
      m->counter_idx[k].push_back(ns + d->n_slot + n_counter++);
  m->scal_bytes = 8 * (ns + d->n_slot + n_counter);
  // tiles and partials
  size_t part_bytes = 0;
  m->part_off.resize(d->n_kernel);
  m->ntiles.resize(d->n_kernel);
  m->tiles.resize(d->n_kernel);
  for (int k = 0; k < d->n_kernel; ++k) {
    for (int l = 0; l < d->kernels[k].n_launch; ++l) {
      const pr_launch_desc& ld = d->kernels[k].launches[l];
      std::array<int, 3> t = {1, 1, 1};
      unsigned nt = 1;
      if (ld.tile[0] > 0) {
        for (int a = 0; a < 3; ++a) {
          int ta = ld.tile[a] > 0 ? ld.tile[a] : 1;
          t[a] = (m->n[a] + ta - 1) / ta;
          nt *= (unsigned)t[a];
        }
      }
      m->tiles[k].push_back(t);
      m->ntiles[k].push_back(nt);
      m->part_off[k].push_back(part_bytes);
      part_bytes += ((size_t)ld.n_reduce * nt * 8 + 255) / 256 * 256;
    }
  }
  hipError_t e;
  e = hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    int rc = hip_fail(m, e, "hipStreamCreateWithFlags");
    delete m;
    return rc;
  }
  m->own_stream = true;
  for (int s = 0; s < ns; ++s) {
    if (d->statics[s].realm != PR_ARRAY) continue;
    size_t bytes = (size_t)m->cells * ctype_size(d->statics[s].ctype);
    for (int b = 0; b < (m->stored_any[s] ? 2 : 1); ++b) {
      e = hipMalloc(&m->buf[b][s], bytes);
      if (e == hipSuccess) e = hipMemsetAsync(m->buf[b][s], 0, bytes, m->stream);
      if (e != hipSuccess) {
        int rc = fail(m, PR_E_NOMEM, "hipMalloc of static buffer failed");
        pr_destroy(m);
        return rc;
      }
    }
  }
  e = hipMalloc(&m->scal, m->scal_bytes);
  if (e == hipSuccess) e = hipMemsetAsync(m->scal, 0, m->scal_bytes, m->stream);
  if (e == hipSuccess) e = hipMalloc(&m->part, part_bytes > 0 ? part_bytes : 256);
  if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
  if (e != hipSuccess) {
    pr_destroy(m);
    return PR_E_NOMEM;
  }
  *out = m;
  return PR_OK;
}

int pr_destroy(pr_machine* m) {
  if (!m) return PR_E_INVALID;
  if (m->stream) hipStreamSynchronize(m->stream);
  for (auto& kv : m->graphs) (void)hipGraphExecDestroy(kv.second);
  for (int b = 0; b < 2; ++b)
    for (void* p : m->buf[b])
      if (p) (void)hipFree(p);
  for (auto& v : m->snaps)
    for (void* p : v)
      if (p) (void)hipFree(p);
  for (char* p : m->snap_scal)
    if (p) (void)hipFree(p);
  if (m->scal) hipFree(m->scal);
  if (m->part) hipFree(m->part);
  if (m->own_stream && m->stream) hipStreamDestroy(m->stream);
  delete m;
  return PR_OK;
}

int pr_set_stream(pr_machine* m, void* stream) {
  if (!m) return PR_E_INVALID;
  PR_HIP(m, hipStreamSynchronize(m->stream));
  if (m->own_stream) hipStreamDestroy(m->stream);
  m->stream = static_cast<hipStream_t>(stream);
  m->own_stream = false;
  for (auto& kv : m->graphs) (void)hipGraphExecDestroy(kv.second);
  m->graphs.clear();
  return PR_OK;
}

int pr_sync(pr_machine* m) {
  if (!m) return PR_E_INVALID;
  PR_HIP(m, hipStreamSynchronize(m->stream));
  return PR_OK;
}

const char* pr_last_error(const pr_machine* m) { return m ? m->err.c_str() : "null handle"; }

int pr_run(pr_machine* m, int k, int count) {
  if (!m || k < 0 || k >= m->d->n_kernel || count < 0) return PR_E_INVALID;
  for (int c = 0; c < count; ++c) {
    int rc = launch_kernel_once(m, k);
    if (rc) return rc;
  }
  return PR_OK;
}

int pr_run_timed(pr_machine* m, int k, int count, double* launch_ms) {
  if (!m || k < 0 || k >= m->d->n_kernel || count < 0 || !launch_ms) return PR_E_INVALID;
  const pr_kernel_desc& kd = m->d->kernels[k];
  const int nl = kd.n_launch;
  std::vector<hipEvent_t> ev((size_t)2 * nl * (count > 0 ? count : 1));
  for (auto& e : ev) PR_HIP(m, hipEventCreate(&e));
  m->timing = ev.data();
  m->timing_stride = nl;
  int rc = PR_OK;
  for (int c = 0; c < count && rc == PR_OK; ++c) {
    m->timing_call = c;
    rc = launch_kernel_once(m, k);
  }
  m->timing = nullptr;
  if (rc == PR_OK) {
    PR_HIP(m, hipStreamSynchronize(m->stream));
    for (int l = 0; l < nl; ++l) launch_ms[l] = 0.0;
    for (int c = 0; c < count; ++c)
      for (int l = 0; l < nl; ++l) {
        float ms = 0.f;
        PR_HIP(m, hipEventElapsedTime(&ms, ev[2 * ((size_t)c * nl + l)], ev[2 * ((size_t)c * nl + l) + 1]));
        launch_ms[l] += ms;
      }
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  return rc;
}

int pr_run_graph(pr_machine* m, int k, int count) {
  if (!m || k < 0 || k >= m->d->n_kernel || count < 1) return PR_E_INVALID;
  GraphKey key{k, count, parity_key(m)};
  auto it = m->graphs.find(key);
  if (it == m->graphs.end()) {
    std::vector<int> saved = m->cur;
    hipGraph_t g;
    PR_HIP(m, hipStreamBeginCapture(m->stream, hipStreamCaptureModeThreadLocal));
    int rc = PR_OK;
    for (int c = 0; c < count && rc == PR_OK; ++c) rc = launch_kernel_once(m, k);
    hipError_t e = hipStreamEndCapture(m->stream, &g);
    m->cur = saved;
    if (rc) return rc;
    if (e != hipSuccess) return hip_fail(m, e, "hipStreamEndCapture");
    hipGraphExec_t ex;
    e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (e != hipSuccess) return hip_fail(m, e, "hipGraphInstantiate");
    it = m->graphs.emplace(key, ex).first;
  }
  PR_HIP(m, hipGraphLaunch(it->second, m->stream));
  const pr_kernel_desc& kd = m->d->kernels[k];
  if (count % 2)
    for (int i = 0; i < kd.n_stored; ++i) m->cur[kd.stored[i]] ^= 1;
  return PR_OK;
}

size_t pr_static_bytes(const pr_machine* m, int s) {
  if (!m || s < 0 || s >= m->d->n_static) return 0;
  const pr_static_desc& sd = m->d->statics[s];
  if (sd.realm == PR_ARRAY) return (size_t)m->cells * ctype_size(sd.ctype);
  return ctype_size(sd.ctype);
}

int pr_static_devptr(pr_machine* m, int s, void** p) {
  if (!m || !p || s < 0 || s >= m->d->n_static) return PR_E_INVALID;
  if (m->d->statics[s].realm == PR_ARRAY)
    *p = m->buf[m->cur[s]][s];
  else
    *p = m->scal + 8 * s;
  return PR_OK;
}

int pr_static_write(pr_machine* m, int s, const void* src, size_t bytes, int src_on_device) {
  if (!m || !src || s < 0 || s >= m->d->n_static) return PR_E_INVALID;
  if (bytes != pr_static_bytes(m, s)) return fail(m, PR_E_INVALID, "static write: size mismatch");
  void* dst;
  pr_static_devptr(m, s, &dst);
  PR_HIP(m, hipMemcpyAsync(dst, src, bytes, src_on_device ? hipMemcpyDeviceToDevice
                                                          : hipMemcpyHostToDevice, m->stream));
  if (!src_on_device) PR_HIP(m, hipStreamSynchronize(m->stream));
  return PR_OK;
}

int pr_static_read(pr_machine* m, int s, void* dst, size_t bytes, int dst_on_device) {
  if (!m || !dst || s < 0 || s >= m->d->n_static) return PR_E_INVALID;
  if (bytes != pr_static_bytes(m, s)) return fail(m, PR_E_INVALID, "static read: size mismatch");
  void* src;
  pr_static_devptr(m, s, &src);
  PR_HIP(m, hipMemcpyAsync(dst, src, bytes, dst_on_device ? hipMemcpyDeviceToDevice
                                                          : hipMemcpyDeviceToHost, m->stream));
  if (!dst_on_device) PR_HIP(m, hipStreamSynchronize(m->stream));
  return PR_OK;
}

int pr_size(const pr_machine* m, int axis) {
  if (!m || axis < 0 || axis >= 3) return 0;
  return m->local[axis];
}

int pr_memory_size(const pr_machine* m, int axis) {
  if (!m || axis < 0 || axis >= 3) return 0;
  return m->n[axis];
}

long long pr_cells(const pr_machine* m) { return m ? m->cells : 0; }

// ---------------------------------------------------------------- whole-machine state files
// Layout: "PRSTATE1", then int32 abi, ndim, local[3], n_static and the machine name
// (uint32 length + bytes); per user static: name, int32 realm, int32 ctype, uint64 bytes,
// data.  Scratch statics are not part of the state.

static bool put(FILE* f, const void* p, size_t n) { return std::fwrite(p, 1, n, f) == n; }
static bool get(FILE* f, void* p, size_t n) { return std::fread(p, 1, n, f) == n; }
static bool put_str(FILE* f, const char* s) {
  const uint32_t n = (uint32_t)std::strlen(s);
  return put(f, &n, 4) && put(f, s, n);
}
static bool get_str(FILE* f, std::string& s) {
  uint32_t n;
  if (!get(f, &n, 4) || n > (1u << 20)) return false;
  s.resize(n);
  return n == 0 || get(f, &s[0], n);
}

int pr_state_save(pr_machine* m, const char* path) {
  if (!m || !path) return PR_E_INVALID;
  FILE* f = std::fopen(path, "wb");
  if (!f) return fail(m, PR_E_INVALID, std::string("state save: cannot open ") + path);
  const pr_machine_desc* d = m->d;
  int32_t head[7] = {PR_ABI_VERSION, d->ndim, m->local[0], m->local[1], m->local[2], 0, 0};
  for (int s = 0; s < d->n_static; ++s)
    if (!(d->statics[s].flags & PR_STATIC_SCRATCH)) ++head[5];
  bool ok = put(f, "PRSTATE1", 8) && put(f, head, 6 * 4) && put_str(f, d->name);
  std::vector<char> host;
  for (int s = 0; ok && s < d->n_static; ++s) {
    const pr_static_desc& sd = d->statics[s];
    if (sd.flags & PR_STATIC_SCRATCH) continue;
    const uint64_t n = pr_static_bytes(m, s);
    host.resize(n);
    if (pr_static_read(m, s, host.data(), n, 0) != PR_OK) {
      std::fclose(f);
      return PR_E_HIP;
    }
    const int32_t rc[2] = {sd.realm, sd.ctype};
    ok = put_str(f, sd.name) && put(f, rc, 8) && put(f, &n, 8) && put(f, host.data(), n);
  }
  ok = (std::fclose(f) == 0) && ok;
  return ok ? PR_OK : fail(m, PR_E_INVALID, std::string("state save: write failed: ") + path);
}

int pr_state_load(pr_machine* m, const char* path) {
  if (!m || !path) return PR_E_INVALID;
  FILE* f = std::fopen(path, "rb");
  if (!f) return fail(m, PR_E_INVALID, std::string("state load: cannot open ") + path);
  const pr_machine_desc* d = m->d;
  auto bad = [&](const std::string& why) {
    std::fclose(f);
    return fail(m, PR_E_INVALID, "state load: " + why);
  };
  char magic[8];
  int32_t head[6];
  std::string name;
  if (!get(f, magic, 8) || std::memcmp(magic, "PRSTATE1", 8) || !get(f, head, 24) ||
      !get_str(f, name))
    return bad("not a state file");
  if (head[0] != PR_ABI_VERSION) return bad("ABI version differs");
  if (name != d->name || head[1] != d->ndim) return bad("saved by machine " + name);
  for (int a = 0; a < 3; ++a)
    if (head[2 + a] != m->local[a]) return bad("grid size differs");
  int n_user = 0;
  for (int s = 0; s < d->n_static; ++s)
    if (!(d->statics[s].flags & PR_STATIC_SCRATCH)) ++n_user;
  if (head[5] != n_user) return bad("static count differs");
  // read everything before writing anything: a mismatch leaves the machine untouched
  std::vector<std::vector<char>> data(d->n_static);
  for (int s = 0; s < d->n_static; ++s) {
    const pr_static_desc& sd = d->statics[s];
    if (sd.flags & PR_STATIC_SCRATCH) continue;
    std::string sname;
    int32_t rc[2];
    uint64_t n;
    if (!get_str(f, sname) || !get(f, rc, 8) || !get(f, &n, 8)) return bad("truncated");
    if (sname != sd.name || rc[0] != sd.realm || rc[1] != sd.ctype ||
        n != pr_static_bytes(m, s))
      return bad("static " + sname + " differs from " + sd.name);
    data[s].resize(n);
    if (!get(f, data[s].data(), n)) return bad("truncated");
  }
  std::fclose(f);
  for (int s = 0; s < d->n_static; ++s) {
    if (data[s].empty()) continue;
    int rc = pr_static_write(m, s, data[s].data(), data[s].size(), 0);
    if (rc) return rc;
  }
  return PR_OK;
}

}  // extern "C"

namespace {

// One launch copies every static of a snapshot (arrays and 8-byte scalar entries): a
// checkpointed adjoint saves or restores a snapshot per time step, and one
// hipMemcpyAsync per static would cost a copy launch each.
constexpr int PR_COPY_MAX = 32;
struct CopyJob {
  char* dst;
  const char* src;
  unsigned long long n;
};
struct CopyBatch {
  CopyJob job[PR_COPY_MAX];
};

__global__ void __launch_bounds__(256) pr_multi_copy(CopyBatch b) {
  const CopyJob j = b.job[blockIdx.y];
  const unsigned long long stride = (unsigned long long)gridDim.x * blockDim.x;
  const unsigned long long t = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
  const bool vec = (((uintptr_t)j.dst | (uintptr_t)j.src) & 15) == 0;
  const unsigned long long n16 = vec ? j.n / 16 : 0;
  for (unsigned long long i = t; i < n16; i += stride)
    reinterpret_cast<uint4*>(j.dst)[i] = reinterpret_cast<const uint4*>(j.src)[i];
  for (unsigned long long i = n16 * 16 + t; i < j.n; i += stride) j.dst[i] = j.src[i];
}

int snapshot_copy(pr_machine* m, int slot, bool save, bool scalars_only = false) {
  CopyBatch b;
  int count = 0;
  unsigned long long most = 0;
  auto flush = [&]() -> int {
    if (!count) return PR_OK;
    unsigned long long blocks = (most / 16 + 255) / 256;
    if (blocks < 1) blocks = 1;
    if (blocks > 2048) blocks = 2048;
    hipLaunchKernelGGL(pr_multi_copy, dim3((unsigned)blocks, (unsigned)count), dim3(256), 0,
                       m->stream, b);
    if (hipPeekAtLastError() != hipSuccess) return fail(m, PR_E_HIP, "snapshot copy launch");
    count = 0;
    most = 0;
    return PR_OK;
  };
  for (int s : m->snap_sel) {
    char* live;
    char* keep;
    unsigned long long n;
    if (scalars_only && m->d->statics[s].realm == PR_ARRAY) continue;
    if (m->d->statics[s].realm == PR_ARRAY) {
      live = (char*)m->buf[m->cur[s]][s];
      keep = (char*)m->snaps[slot][s];
      n = pr_static_bytes(m, s);
    } else {
      live = m->scal + 8 * s;
      keep = m->snap_scal[slot] + 8 * s;
      n = 8;
    }
    b.job[count].dst = save ? keep : live;
    b.job[count].src = save ? live : keep;
    b.job[count].n = n;
    if (n > most) most = n;
    if (++count == PR_COPY_MAX) {
      int rc = flush();
      if (rc) return rc;
    }
  }
  return flush();
}

}  // namespace

extern "C" {

int pr_snapshot_alloc(pr_machine* m, int n_slots, const int* statics, int n_statics) {
  if (!m || n_slots < 0 || n_statics < 0 || (n_statics > 0 && !statics)) return PR_E_INVALID;
  const int ns = m->d->n_static;
  std::vector<int> sel(statics, statics + n_statics);
  for (int s : sel)
    if (s < 0 || s >= ns) return fail(m, PR_E_INVALID, "snapshot static index out of range");
  if (!m->snaps.empty() && sel != m->snap_sel) {
    // a different selection: drop the old snapshots
    for (auto& v : m->snaps)
      for (void* p : v)
        if (p) (void)hipFree(p);
    for (char* p : m->snap_scal)
      if (p) (void)hipFree(p);
    m->snaps.clear();
    m->snap_scal.clear();
  }
  m->snap_sel = sel;
  while ((int)m->snaps.size() < n_slots) {
    std::vector<void*> v(ns, nullptr);
    for (int s : sel) {
      if (m->d->statics[s].realm != PR_ARRAY) continue;
      if (hipMalloc(&v[s], pr_static_bytes(m, s)) != hipSuccess) {
        for (void* p : v)
          if (p) (void)hipFree(p);
        return fail(m, PR_E_NOMEM, "snapshot allocation failed");
      }
    }
    char* sc = nullptr;
    if (hipMalloc(&sc, 8 * (size_t)ns) != hipSuccess) return fail(m, PR_E_NOMEM, "snapshot scalars");
    m->snaps.push_back(v);
    m->snap_scal.push_back(sc);
  }
  return PR_OK;
}

int pr_snapshot_save(pr_machine* m, int slot) {
  if (!m || slot < 0 || slot >= (int)m->snaps.size()) return PR_E_INVALID;
  return snapshot_copy(m, slot, /*save=*/true);
}

int pr_snapshot_load(pr_machine* m, int slot) {
  if (!m || slot < 0 || slot >= (int)m->snaps.size()) return PR_E_INVALID;
  return snapshot_copy(m, slot, /*save=*/false);
}

int pr_adjoint_forward_seq(pr_machine* m, const int* fwd, int n_fwd, int nsteps, int every) {
  if (!m || !fwd || n_fwd < 1 || nsteps < 0 || every < 1) return PR_E_INVALID;
  for (int i = 0; i < n_fwd; ++i)
    if (fwd[i] < 0 || fwd[i] >= m->d->n_kernel)
      return fail(m, PR_E_INVALID, "adjoint: kernel index out of range");
  const int S = (nsteps + every - 1) / every;
  if ((int)m->snaps.size() < S + every)
    return fail(m, PR_E_INVALID, "pr_snapshot_alloc needs ceil(n/every) + every slots");
  for (int n = 0; n < nsteps; ++n) {
    int rc;
    if (n % every == 0 && (rc = pr_snapshot_save(m, n / every))) return rc;
    for (int i = 0; i < n_fwd; ++i)
      if ((rc = launch_kernel_once(m, fwd[i]))) return rc;
  }
  return PR_OK;
}

int pr_adjoint_forward(pr_machine* m, int fwd, int nsteps, int every) {
  return pr_adjoint_forward_seq(m, &fwd, 1, nsteps, every);
}

namespace {

bool lists(const int* xs, int n, int s) {
  for (int i = 0; i < n; ++i)
    if (xs[i] == s) return true;
  return false;
}

}  // namespace

int pr_adjoint_backward_seq(pr_machine* m, const int* fwd, int n_fwd, int tape_after,
                            const int* adj, int n_adj, int nsteps, int every) {
  if (!m || !fwd || !adj || n_fwd < 1 || n_adj < 1 || nsteps < 0 || every < 1 ||
      tape_after < 0 || tape_after >= n_fwd)
    return PR_E_INVALID;
  const pr_machine_desc* d = m->d;
  for (int i = 0; i < n_fwd; ++i)
    if (fwd[i] < 0 || fwd[i] >= d->n_kernel)
      return fail(m, PR_E_INVALID, "adjoint: kernel index out of range");
  for (int i = 0; i < n_adj; ++i)
    if (adj[i] < 0 || adj[i] >= d->n_kernel)
      return fail(m, PR_E_INVALID, "adjoint: kernel index out of range");
  const int S = (nsteps + every - 1) / every;
  if ((int)m->snaps.size() < S + every)
    return fail(m, PR_E_INVALID, "pr_snapshot_alloc needs ceil(n/every) + every slots");
  const int ns = d->n_static;
  std::vector<char> snap(ns, 0), fwd_writes(ns, 0), pre_writes(ns, 0);
  for (int s : m->snap_sel) snap[s] = 1;
  for (int i = 0; i < n_fwd; ++i) {
    const pr_kernel_desc& kd = d->kernels[fwd[i]];
    for (int w = 0; w < kd.n_written; ++w) {
      fwd_writes[kd.written[w]] = 1;
      if (i < tape_after) pre_writes[kd.written[w]] = 1;
    }
  }
  // the adjoint kernels read the state at the tape point; a linear scheme's adjoint reads
  // none of what the forward writes, and then needs no replay
  bool needs_state = false, needs_scalars = false;
  for (int i = 0; i < n_adj; ++i) {
    const pr_kernel_desc& ka = d->kernels[adj[i]];
    for (int r = 0; r < ka.n_loaded; ++r) {
      const int s = ka.loaded[r];
      if (!snap[s] || !fwd_writes[s]) continue;
      needs_state = true;
      if (d->statics[s].realm != PR_ARRAY) needs_scalars = true;
    }
    for (int w = 0; w < ka.n_written; ++w)
      if (snap[ka.written[w]])
        return fail(m, PR_E_UNSUPPORTED, std::string("adjoint kernel ") + ka.name +
                                             " writes a snapshotted static");
  }
  int rc;
  if (!needs_state) {
    for (int j = 0; j < nsteps; ++j)
      for (int i = 0; i < n_adj; ++i)
        if ((rc = launch_kernel_once(m, adj[i]))) return rc;
    return S > 0 ? pr_snapshot_load(m, 0) : PR_OK;
  }
  // A kernel before the tape point writes its arrays into the step's own tape slot, so it
  // must not read what it overwrites there.
  for (int i = 0; i < tape_after; ++i) {
    const pr_kernel_desc& kd = d->kernels[fwd[i]];
    for (int w = 0; w < kd.n_written; ++w) {
      const int s = kd.written[w];
      if (!snap[s] || d->statics[s].realm != PR_ARRAY) continue;
      for (int i2 = 0; i2 <= i; ++i2) {
        const pr_kernel_desc& k2 = d->kernels[fwd[i2]];
        if (lists(k2.loaded, k2.n_loaded, s))
          return fail(m, PR_E_UNSUPPORTED, std::string("kernel ") + kd.name +
                                               " before the tape point rewrites a static it reads");
      }
    }
  }
  // Replay by rotation: the state of step j lives in tape slot j (the segment's checkpoint
  // for j = 0, snapshot slot S + j after), each kernel reads it there and writes what it
  // stores straight into the slot of the step the value belongs to, so no state is copied
  // per step.  Arrays the forward never writes stay in the checkpoint.
  std::vector<const void*> cur_ov(ns, nullptr);
  std::vector<void*> next_ov(ns, nullptr);
  auto slot_of = [&](int seg, int j) { return j == 0 ? seg : S + j; };
  for (int seg = S - 1; seg >= 0; --seg) {
    const int start = seg * every;
    const int len = (nsteps - start < every) ? nsteps - start : every;
    // scalars evolve through the replay exactly as in the forward pass
    if ((rc = snapshot_copy(m, seg, /*save=*/false, /*scalars_only=*/true))) return rc;
    for (int j = 0; j < len; ++j) {
      std::vector<void*> here(ns, nullptr), ahead(ns, nullptr);
      for (int s = 0; s < ns; ++s) {
        if (!snap[s] || d->statics[s].realm != PR_ARRAY) continue;
        here[s] = m->snaps[fwd_writes[s] ? slot_of(seg, j) : seg][s];
        ahead[s] = (j + 1 < len) ? m->snaps[slot_of(seg, j + 1)][s] : nullptr;
        cur_ov[s] = here[s];
      }
      for (int i = 0; i < n_fwd; ++i) {
        if (i == tape_after && needs_scalars &&
            (rc = snapshot_copy(m, S + j, /*save=*/true, /*scalars_only=*/true)))
          return rc;
        if (i >= tape_after && j + 1 == len) break;   // the next checkpoint holds the rest
        const pr_kernel_desc& kd = d->kernels[fwd[i]];
        for (int w = 0; w < kd.n_stored; ++w) {
          const int s = kd.stored[w];
          if (snap[s]) next_ov[s] = (i < tape_after) ? here[s] : ahead[s];
        }
        if ((rc = launch_kernel_ov(m, fwd[i], cur_ov.data(), next_ov.data()))) return rc;
        for (int w = 0; w < kd.n_stored; ++w) {
          const int s = kd.stored[w];
          if (snap[s]) cur_ov[s] = next_ov[s];
        }
      }
    }
    for (int j = len - 1; j >= 0; --j) {
      for (int s = 0; s < ns; ++s)
        if (snap[s] && d->statics[s].realm == PR_ARRAY)
          cur_ov[s] = m->snaps[fwd_writes[s] ? slot_of(seg, j) : seg][s];
      if (needs_scalars && (rc = snapshot_copy(m, S + j, /*save=*/false, /*scalars_only=*/true)))
        return rc;
      for (int i = 0; i < n_adj; ++i)
        if ((rc = launch_kernel_ov(m, adj[i], cur_ov.data(), nullptr))) return rc;
    }
  }
  return S > 0 ? pr_snapshot_load(m, 0) : PR_OK;
}

int pr_adjoint_backward(pr_machine* m, int fwd, int adj, int nsteps, int every) {
  return pr_adjoint_backward_seq(m, &fwd, 1, 0, &adj, 1, nsteps, every);
}


// ---------------------------------------------------------------- batches
}  // extern "C"

struct pr_batch {
  const pr_machine_desc* d = nullptr;
  std::vector<pr_machine*> grids;
  hipStream_t stream = nullptr;
  long long cells = 0;
  // per (kernel, launch): grid table
  std::vector<std::vector<int*>> tile_grid;
  std::vector<std::vector<unsigned*>> tile_base;
  std::vector<std::vector<unsigned>> total;
  // per (kernel, launch, parity of every grid): packed argument blocks on the device
  std::map<std::tuple<int, int, std::vector<char>>, char*> args;
  std::map<GraphKey, hipGraphExec_t> graphs;
  std::string err;
};

namespace {

int bfail(pr_batch* b, int code, const std::string& msg) {
  if (b) b->err = msg;
  return code;
}

// Parity of every static of every grid: grids can also be run one at a time
// (pr_batch_grid + pr_run), so their parities need not agree.
std::vector<char> batch_parity_key(const pr_batch* b) {
  std::vector<char> p;
  for (const pr_machine* m : b->grids)
    for (int c : m->cur) p.push_back((char)(c & 1));
  return p;
}

// Build the launch context of grid machine m for launch (k, l), as launch_kernel_once does.
void grid_ctx(pr_machine* m, int k, int l, std::vector<const void*>& cur, std::vector<void*>& next,
              pr_launch_ctx& ctx) {
  const int ns = m->d->n_static;
  cur.assign(ns, nullptr);
  next.assign(ns, nullptr);
  for (int s = 0; s < ns; ++s) {
    if (m->d->statics[s].realm != PR_ARRAY) continue;
    cur[s] = m->buf[m->cur[s]][s];
    if (m->stored_any[s]) next[s] = m->buf[1 - m->cur[s]][s];
  }
  ctx.cur = cur.data();
  ctx.next = next.data();
  ctx.scal = m->scal;
  ctx.part = m->part + m->part_off[k][l];
  ctx.counter = reinterpret_cast<unsigned*>(m->scal + 8 * m->counter_idx[k][l]);
  for (int a = 0; a < 3; ++a) {
    ctx.n[a] = m->n[a];
    ctx.tiles[a] = m->tiles[k][l][a];
  }
  ctx.ntiles = m->ntiles[k][l];
  ctx.stream = m->stream;
}

// Device block of packed arguments for (k, l) at the grids' current parity; built on first use
// with synchronous copies, so a run that is captured must have built it beforehand.
int batch_args(pr_batch* b, int k, int l, char** out) {
  const auto key = std::make_tuple(k, l, batch_parity_key(b));
  auto it = b->args.find(key);
  if (it != b->args.end()) {
    *out = it->second;
    return PR_OK;
  }
  const pr_launch_desc& ld = b->d->kernels[k].launches[l];
  const size_t sz = ld.args_size;
  std::vector<char> host(sz * b->grids.size());
  std::vector<const void*> cur;
  std::vector<void*> next;
  for (size_t g = 0; g < b->grids.size(); ++g) {
    pr_launch_ctx ctx;
    grid_ctx(b->grids[g], k, l, cur, next, ctx);
    ld.pack(&ctx, host.data() + g * sz);
  }
  char* dev = nullptr;
  if (hipMalloc(&dev, host.size()) != hipSuccess) return bfail(b, PR_E_NOMEM, "batch args");
  if (hipMemcpy(dev, host.data(), host.size(), hipMemcpyHostToDevice) != hipSuccess)
    return bfail(b, PR_E_HIP, "batch args copy");
  b->args[key] = dev;
  *out = dev;
  return PR_OK;
}

int batch_once(pr_batch* b, int k, double* launch_ms, std::vector<hipEvent_t>* ev) {
  const pr_kernel_desc& kd = b->d->kernels[k];
  for (int l = 0; l < kd.n_launch; ++l) {
    char* args = nullptr;
    int rc = batch_args(b, k, l, &args);
    if (rc) return rc;
    pr_batch_ctx bc;
    bc.args = args;
    bc.tile_grid = b->tile_grid[k][l];
    bc.tile_base = b->tile_base[k][l];
    bc.total_tiles = b->total[k][l];
    bc.n_grids = (int)b->grids.size();
    bc.stream = b->stream;
    if (ev) (void)hipEventRecord((*ev)[2 * l], b->stream);
    if (kd.launches[l].batched(&bc) != 0)
      return bfail(b, PR_E_HIP, std::string("batched launch: ") + hipGetErrorString(hipGetLastError()));
    if (ev) (void)hipEventRecord((*ev)[2 * l + 1], b->stream);
  }
  for (pr_machine* m : b->grids)
    for (int i = 0; i < kd.n_stored; ++i) m->cur[kd.stored[i]] ^= 1;
  return PR_OK;
}

}  // namespace

extern "C" {

int pr_batch_create(const pr_machine_desc* d, int n_grids, const int* sizes, pr_batch** out) {
  if (!d || n_grids < 1 || !sizes || !out) return PR_E_INVALID;
  if (d->abi != PR_ABI_VERSION) return PR_E_ABI;
  *out = nullptr;
  for (int k = 0; k < d->n_kernel; ++k)
    for (int l = 0; l < d->kernels[k].n_launch; ++l)
      if (!d->kernels[k].launches[l].batched || !d->kernels[k].launches[l].pack)
        return PR_E_UNSUPPORTED;   // generated without Setup.batched
  pr_batch* b = new pr_batch();
  b->d = d;
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
    delete b;
    return PR_E_HIP;
  }
  for (int g = 0; g < n_grids; ++g) {
    pr_machine* m = nullptr;
    int rc = pr_create(d, sizes + 3 * g, &m);
    if (rc == PR_OK) rc = pr_set_stream(m, b->stream);
    if (rc) {
      b->err = "grid " + std::to_string(g) + ": " + (m ? m->err : std::string("pr_create"));
      pr_batch_destroy(b);
      return rc;
    }
    b->grids.push_back(m);
    b->cells += m->cells;
  }
  b->tile_grid.resize(d->n_kernel);
  b->tile_base.resize(d->n_kernel);
  b->total.resize(d->n_kernel);
  for (int k = 0; k < d->n_kernel; ++k) {
    for (int l = 0; l < d->kernels[k].n_launch; ++l) {
      std::vector<int> tg;
      std::vector<unsigned> tb;
      for (int g = 0; g < n_grids; ++g) {
        tb.push_back((unsigned)tg.size());
        for (unsigned t = 0; t < b->grids[g]->ntiles[k][l]; ++t) tg.push_back(g);
      }
      int* dtg = nullptr;
      unsigned* dtb = nullptr;
      if (hipMalloc(&dtg, tg.size() * sizeof(int)) != hipSuccess ||
          hipMalloc(&dtb, tb.size() * sizeof(unsigned)) != hipSuccess ||
          hipMemcpy(dtg, tg.data(), tg.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
          hipMemcpy(dtb, tb.data(), tb.size() * sizeof(unsigned), hipMemcpyHostToDevice) != hipSuccess) {
        pr_batch_destroy(b);
        return PR_E_NOMEM;
      }
      b->tile_grid[k].push_back(dtg);
      b->tile_base[k].push_back(dtb);
      b->total[k].push_back((unsigned)tg.size());
    }
  }
  *out = b;
  return PR_OK;
}

int pr_batch_destroy(pr_batch* b) {
  if (!b) return PR_E_INVALID;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  for (auto& kv : b->graphs) (void)hipGraphExecDestroy(kv.second);
  for (auto& kv : b->args) (void)hipFree(kv.second);
  for (auto& v : b->tile_grid)
    for (int* p : v) (void)hipFree(p);
  for (auto& v : b->tile_base)
    for (unsigned* p : v) (void)hipFree(p);
  for (pr_machine* m : b->grids) pr_destroy(m);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return PR_OK;
}

pr_machine* pr_batch_grid(pr_batch* b, int g) {
  if (!b || g < 0 || g >= (int)b->grids.size()) return nullptr;
  return b->grids[g];
}

int pr_batch_run(pr_batch* b, int k, int count) {
  if (!b || k < 0 || k >= b->d->n_kernel || count < 0) return PR_E_INVALID;
  for (int c = 0; c < count; ++c) {
    int rc = batch_once(b, k, nullptr, nullptr);
    if (rc) return rc;
  }
  return PR_OK;
}

int pr_batch_run_timed(pr_batch* b, int k, int count, double* launch_ms) {
  if (!b || k < 0 || k >= b->d->n_kernel || count < 0 || !launch_ms) return PR_E_INVALID;
  const int nl = b->d->kernels[k].n_launch;
  std::vector<std::vector<hipEvent_t>> evs(count, std::vector<hipEvent_t>(2 * nl));
  for (auto& v : evs)
    for (auto& e : v)
      if (hipEventCreate(&e) != hipSuccess) return bfail(b, PR_E_HIP, "hipEventCreate");
  int rc = PR_OK;
  for (int c = 0; c < count && rc == PR_OK; ++c) rc = batch_once(b, k, nullptr, &evs[c]);
  if (rc == PR_OK) {
    (void)hipStreamSynchronize(b->stream);
    for (int l = 0; l < nl; ++l) launch_ms[l] = 0;
    for (int c = 0; c < count; ++c)
      for (int l = 0; l < nl; ++l) {
        float ms = 0;
        (void)hipEventElapsedTime(&ms, evs[c][2 * l], evs[c][2 * l + 1]);
        launch_ms[l] += ms;
      }
  }
  for (auto& v : evs)
    for (auto& e : v) (void)hipEventDestroy(e);
  return rc;
}

int pr_batch_run_graph(pr_batch* b, int k, int count) {
  if (!b || k < 0 || k >= b->d->n_kernel || count < 1) return PR_E_INVALID;
  GraphKey key{k, count, batch_parity_key(b)};
  auto it = b->graphs.find(key);
  if (it == b->graphs.end()) {
    // build every argument block the captured sequence needs before capturing
    std::vector<std::vector<int>> saved;
    for (pr_machine* m : b->grids) saved.push_back(m->cur);
    const pr_kernel_desc& kd = b->d->kernels[k];
    for (int c = 0; c < count && c < 2; ++c) {
      for (int l = 0; l < kd.n_launch; ++l) {
        char* a;
        int rc = batch_args(b, k, l, &a);
        if (rc) return rc;
      }
      for (pr_machine* m : b->grids)
        for (int i = 0; i < kd.n_stored; ++i) m->cur[kd.stored[i]] ^= 1;
    }
    for (size_t g = 0; g < b->grids.size(); ++g) b->grids[g]->cur = saved[g];
    hipGraph_t gr;
    if (hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
      return bfail(b, PR_E_HIP, "begin capture");
    int rc = PR_OK;
    for (int c = 0; c < count && rc == PR_OK; ++c) rc = batch_once(b, k, nullptr, nullptr);
    hipError_t e = hipStreamEndCapture(b->stream, &gr);
    for (size_t g = 0; g < b->grids.size(); ++g) b->grids[g]->cur = saved[g];
    if (rc) return rc;
    if (e != hipSuccess) return bfail(b, PR_E_HIP, "end capture");
    hipGraphExec_t ex;
    e = hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0);
    (void)hipGraphDestroy(gr);
    if (e != hipSuccess) return bfail(b, PR_E_HIP, "graph instantiate");
    it = b->graphs.emplace(key, ex).first;
  }
  if (hipGraphLaunch(it->second, b->stream) != hipSuccess) return bfail(b, PR_E_HIP, "graph launch");
  const pr_kernel_desc& kd = b->d->kernels[k];
  if (count % 2)
    for (pr_machine* m : b->grids)
      for (int i = 0; i < kd.n_stored; ++i) m->cur[kd.stored[i]] ^= 1;
  return PR_OK;
}

int pr_batch_sync(pr_batch* b) {
  if (!b) return PR_E_INVALID;
  return hipStreamSynchronize(b->stream) == hipSuccess ? PR_OK : PR_E_HIP;
}

const char* pr_batch_last_error(const pr_batch* b) { return b ? b->err.c_str() : "null handle"; }

long long pr_batch_cells(const pr_batch* b) { return b ? b->cells : 0; }

}  // extern "C"

// ---------------------------------------------------------------- running without allocating
// pr_run_noalloc runs a kernel on buffers the caller owns.  What pr_create would allocate
// besides the statics -- the scalar block (scalar statics, scalar slots the finish kernels
// write, per-launch counters), the reduction partials and the scratch arrays of cut
// launches -- lives in `scratch`, a fixed array every GPU holds for the life of the process.
// Host code must never pass `scratch` itself to a kernel or a copy: its name is a host-side
// stand-in, not the device address, which hipGetSymbolAddress gives per device.

static __device__ __align__(256) char scratch[PR_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int PR_NOALLOC_MAX_DEVICES = 64;
constexpr int PR_NOALLOC_MAX_STATICS = 256;
constexpr int PR_NOALLOC_MAX_LAUNCHES = 32;
std::mutex noalloc_mu;
char* noalloc_scratch[PR_NOALLOC_MAX_DEVICES];   // device address of `scratch`, per device

int noalloc_fail(char* err, size_t len, int code, const char* fmt, ...) {
  if (err && len) {
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(err, len, fmt, ap);
    va_end(ap);
  }
  return code;
}

size_t round256(size_t x) { return (x + 255) / 256 * 256; }

}  // namespace

extern "C" int pr_run_noalloc(const pr_machine_desc* d, const int* local_size, int k, int count,
                              void* const* buffers, int* parity, void* scalars, void* stream,
                              char* err, size_t err_len) {
  if (err && err_len) err[0] = 0;
  if (!d || !local_size || !buffers || !parity || count < 0)
    return noalloc_fail(err, err_len, PR_E_INVALID, "null argument or negative count");
  if (d->abi != PR_ABI_VERSION) return noalloc_fail(err, err_len, PR_E_ABI, "descriptor ABI");
  if (k < 0 || k >= d->n_kernel)
    return noalloc_fail(err, err_len, PR_E_INVALID, "kernel index %d out of range", k);
  const int ns = d->n_static;
  const pr_kernel_desc& kd = d->kernels[k];
  if (ns > PR_NOALLOC_MAX_STATICS || kd.n_launch > PR_NOALLOC_MAX_LAUNCHES)
    return noalloc_fail(err, err_len, PR_E_UNSUPPORTED, "machine too large for pr_run_noalloc");
  // grid extents, as pr_create
  int n[3] = {1, 1, 1};
  long long cells = 1;
  for (int a = 0; a < d->ndim; ++a) {
    if (local_size[a] < 1 || local_size[a] < d->halo)
      return noalloc_fail(err, err_len, PR_E_INVALID,
                          "grid extent %d on axis %d is below the machine's halo %d",
                          local_size[a], a, d->halo);
    n[a] = local_size[a] + d->lower_margin[a] + d->upper_margin[a];
    cells *= n[a];
  }
  if (cells >= (1LL << 31))
    return noalloc_fail(err, err_len, PR_E_INVALID, "grid of %lld cells exceeds 2^31 - 1", cells);
  // scratch layout: scalar block, this kernel's partials, scratch arrays
  size_t n_counter = 0, counter0 = 0;
  for (int kk = 0; kk < d->n_kernel; ++kk) {
    if (kk == k) counter0 = n_counter;
    n_counter += (size_t)d->kernels[kk].n_launch;
  }
  const size_t scal_bytes = round256(8 * ((size_t)ns + d->n_slot + n_counter));
  size_t need = scal_bytes;
  size_t part_off[PR_NOALLOC_MAX_LAUNCHES];
  unsigned ntiles[PR_NOALLOC_MAX_LAUNCHES];
  int tiles[PR_NOALLOC_MAX_LAUNCHES][3];
  for (int l = 0; l < kd.n_launch; ++l) {
    const pr_launch_desc& ld = kd.launches[l];
    unsigned nt = 1;
    for (int a = 0; a < 3; ++a) {
      const int ta = ld.tile[a] > 0 ? ld.tile[a] : 1;
      tiles[l][a] = ld.tile[0] > 0 ? (n[a] + ta - 1) / ta : 1;
      nt *= (unsigned)tiles[l][a];
    }
    ntiles[l] = nt;
    part_off[l] = need;
    need += round256((size_t)ld.n_reduce * nt * 8);
  }
  size_t array_off[PR_NOALLOC_MAX_STATICS];
  for (int s = 0; s < ns; ++s) {
    const pr_static_desc& sd = d->statics[s];
    array_off[s] = 0;
    if (sd.realm == PR_ARRAY && (sd.flags & PR_STATIC_SCRATCH)) {
      array_off[s] = need;
      need += round256((size_t)cells * ctype_size(sd.ctype));
    }
  }
  if (need > (size_t)PR_NOALLOC_SCRATCH_BYTES)
    return noalloc_fail(err, err_len, PR_E_INVALID,
                        "kernel %s on this grid needs %zu bytes of scratch; pr_run_noalloc "
                        "holds %zu (PR_NOALLOC_SCRATCH_BYTES): use pr_create and pr_run",
                        kd.name, need, (size_t)PR_NOALLOC_SCRATCH_BYTES);
  // the caller's buffers: one per array static, a second for those this kernel stores
  bool stored[PR_NOALLOC_MAX_STATICS] = {};
  for (int i = 0; i < kd.n_stored; ++i) stored[kd.stored[i]] = true;
  bool any_scalar = false;
  for (int s = 0; s < ns; ++s) {
    const pr_static_desc& sd = d->statics[s];
    if (sd.realm != PR_ARRAY) {
      any_scalar = true;
      continue;
    }
    if (sd.flags & PR_STATIC_SCRATCH) continue;
    if (parity[s] != 0 && parity[s] != 1)
      return noalloc_fail(err, err_len, PR_E_INVALID, "parity of %s is not 0 or 1", sd.name);
    if (!buffers[2 * s + parity[s]])
      return noalloc_fail(err, err_len, PR_E_INVALID, "no buffer holds %s", sd.name);
    if (stored[s] && !buffers[2 * s + 1 - parity[s]])
      return noalloc_fail(err, err_len, PR_E_INVALID,
                          "kernel %s stores %s: it needs a second buffer", kd.name, sd.name);
  }
  if (any_scalar && !scalars)
    return noalloc_fail(err, err_len, PR_E_INVALID, "the machine has scalar statics: pass them");
  // this device's copy of `scratch`
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess)
    return noalloc_fail(err, err_len, PR_E_HIP, "hipGetDevice: %s", hipGetErrorString(e));
  if (dev < 0 || dev >= PR_NOALLOC_MAX_DEVICES)
    return noalloc_fail(err, err_len, PR_E_UNSUPPORTED, "device %d", dev);
  char* base = nullptr;
  {
    std::lock_guard<std::mutex> lock(noalloc_mu);
    if (!noalloc_scratch[dev]) {
      void* p = nullptr;
      e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
      if (e != hipSuccess)
        return noalloc_fail(err, err_len, PR_E_HIP, "hipGetSymbolAddress: %s", hipGetErrorString(e));
      noalloc_scratch[dev] = static_cast<char*>(p);
    }
    base = noalloc_scratch[dev];
  }
  hipStream_t st = static_cast<hipStream_t>(stream);
  char* scal = base;
  if (any_scalar) {
    e = hipMemcpyAsync(scal, scalars, 8 * (size_t)ns, hipMemcpyDeviceToDevice, st);
    if (e != hipSuccess)
      return noalloc_fail(err, err_len, PR_E_HIP, "scalars in: %s", hipGetErrorString(e));
  }
  const void* cur[PR_NOALLOC_MAX_STATICS];
  void* next[PR_NOALLOC_MAX_STATICS];
  for (int c = 0; c < count; ++c) {
    for (int s = 0; s < ns; ++s) {
      const pr_static_desc& sd = d->statics[s];
      cur[s] = nullptr;
      next[s] = nullptr;
      if (sd.realm != PR_ARRAY) continue;
      if (sd.flags & PR_STATIC_SCRATCH) {
        cur[s] = base + array_off[s];        // cut launches write these through cur
        continue;
      }
      cur[s] = buffers[2 * s + parity[s]];
      if (stored[s]) next[s] = buffers[2 * s + 1 - parity[s]];
    }
    for (int l = 0; l < kd.n_launch; ++l) {
      pr_launch_ctx ctx;
      ctx.cur = cur;
      ctx.next = next;
      ctx.scal = scal;
      ctx.part = base + part_off[l];
      ctx.counter = reinterpret_cast<unsigned*>(scal + 8 * ((size_t)ns + d->n_slot + counter0 + l));
      for (int a = 0; a < 3; ++a) {
        ctx.n[a] = n[a];
        ctx.tiles[a] = tiles[l][a];
      }
      ctx.ntiles = ntiles[l];
      ctx.stream = st;
      if (kd.launches[l].fn(&ctx) != 0)
        return noalloc_fail(err, err_len, PR_E_HIP, "launch %s[%d]: %s", kd.name, l,
                            hipGetErrorString(hipGetLastError()));
    }
    for (int i = 0; i < kd.n_stored; ++i) parity[kd.stored[i]] ^= 1;
  }
  if (any_scalar) {
    e = hipMemcpyAsync(scalars, scal, 8 * (size_t)ns, hipMemcpyDeviceToDevice, st);
    if (e != hipSuccess)
      return noalloc_fail(err, err_len, PR_E_HIP, "scalars out: %s", hipGetErrorString(e));
  }
  return PR_OK;
}
