// Host driver and C ABI of the gfx950 elastic propagator (see include/elastic_iso_3d.h).
//
// One handle = one grid/time axis on one device: the reference's
// init*/allocate*/prop*/deallocate* functions and their file-scope state
// (elastic_iso_float/include/varElaDeclare_3D.h:51-86) become fields of the
// handle. The time loops follow the reference's orders:
//   nonlinear forward  nonlinearPropElasticGpuFunctions_3D.cu:909-925
//   nonlinear adjoint  nonlinearPropElasticGpuFunctions_3D.cu:1330-1345
//   Born forward       BornElasticGpuFunctions_3D.cu:1268-1407
//   Born adjoint       BornElasticGpuFunctions_3D.cu:2147-2281
// with the sponge fused into the stencil launches and the coarse-time
// wavefield of the Born operators kept resident in HBM instead of pinned host
// memory (one HBM3E device holds 288 GB).
//
// Device layout. The caller's arrays are [ny][nx][nz]. Inside the library
// every 3-D array (wavefields, medium, perturbations, coarse wavefields) is
// stored with a row pitch nzi that is a whole number of 128-byte cache lines
// and with z = 0 at element zoff, where zoff + FAT is a line boundary: the
// 64-cell z tiles of the stencil then read whole lines. Arrays are converted
// at the boundary (2-D copies), point lists by a small kernel.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>
#include <tuple>
#include <type_traits>
#include <vector>

#include "../../include/elastic_iso_3d.h"
#include "ela_kernels.hpp"

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIPCHK(expr)                                                                     \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess)                                                                \
      return fail(ELA_ERR_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_),    \
                  __FILE__, __LINE__);                                                   \
  } while (0)

#define CHK(expr)                \
  do {                           \
    int r_ = (expr);             \
    if (r_ != ELA_OK) return r_; \
  } while (0)

const double kC[4] = {1.196289062541883, -0.079752604188901, 0.009570312506634, -6.975446437140719e-04};
const int kCompGrid[9] = {1, 2, 3, 0, 0, 0, 4, 5, 6};

// Device buffer owned by the library.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  int alloc(size_t b) {
    if (b <= bytes && p) return ELA_OK;
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    if (b == 0) return ELA_OK;
    hipError_t e = hipMalloc(&p, b);
    if (e != hipSuccess) return fail(ELA_ERR_MEM, "hipMalloc(%zu) failed: %s", b, hipGetErrorString(e));
    bytes = b;
    return ELA_OK;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  ~DevBuf() { release(); }
};

int64_t round_up(int64_t v, int64_t m) { return (v + m - 1) / m * m; }

// An executable graph kept across calls together with the signature of the
// launches it was captured from (kernel, launch shape and argument bytes of
// every launch, in order).
struct GraphCache {
  std::vector<unsigned char> sig;
  hipGraphExec_t exec = nullptr;
  hipEvent_t done = nullptr;   // recorded on the stream after the exec's last replay
};

// The handle's graph caches.
enum { G_NL = 0,      // one coarse step of the nonlinear forward
       G_NL_ADJ,      // one coarse step of the nonlinear adjoint
       G_FWD,         // two gradient forward steps (no recording)
       G_FWD_REC,     // two gradient forward steps with receiver recording
       G_STORE,       // a checkpoint segment of forward steps storing their derivative combinations
       G_ADJ,         // a checkpoint segment of transposed steps with imaging
       G_COUNT };

}  // namespace

// Device scratch of the allocation-free entry point (ela_nl_forward_noalloc):
// converted point lists, the position error flag and the time-major receiver
// buffer live here instead of in buffers the library allocates. Every GPU the
// library runs on holds its own copy until the process ends. In host code the
// name `ela_scratch` is a host-side stand-in: its device address on the
// current GPU comes from hipGetSymbolAddress (scratch_address below), and
// only that address is ever passed to a kernel or a copy.
constexpr size_t kScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char ela_scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};   // device address of ela_scratch, per device (null: not resolved)

// Resolve (once per device, with that device current) the device address of
// ela_scratch.
int scratch_address(int dev, void** out) {
  if (dev < 0 || dev >= kMaxDevices) return fail(ELA_ERR_ARG, "device %d beyond the scratch table", dev);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(ela_scratch));
    if (e != hipSuccess || !p)
      return fail(ELA_ERR_HIP, "hipGetSymbolAddress(ela_scratch) failed: %s", hipGetErrorString(e));
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return ELA_OK;
}

// ELA_MEM_DEVICE calls: every caller pointer must be memory the GPU can
// access (device, managed or pinned host memory). A pageable host pointer is
// an argument error, never a kernel that reads host addresses.
int check_dev_ptr(const void* p, const char* what) {
  if (!p) return ELA_OK;
  hipPointerAttribute_t attr;
  std::memset(&attr, 0, sizeof(attr));
  const hipError_t e = hipPointerGetAttributes(&attr, p);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return fail(ELA_ERR_ARG, "%s is not GPU-accessible memory (device-memory call)", what);
  }
  if (attr.type != hipMemoryTypeDevice && attr.type != hipMemoryTypeManaged && attr.type != hipMemoryTypeHost)
    return fail(ELA_ERR_ARG, "%s is not GPU-accessible memory (device-memory call)", what);
  return ELA_OK;
}

}  // namespace

struct ela_handle {
  ela_config cfg;
  int64_t n_model = 0;      // nz*nx*ny (reference layout)
  int64_t nzi = 0;          // internal row pitch
  int64_t zoff = 0;         // internal z origin
  int64_t n_int = 0;        // ny*nx*nzi
  int64_t stride = 0;       // per-array stride in the library's buffers (n_int + guard, rounded up)
  int esize = 4;
  hipStream_t own_stream = nullptr;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  DevBuf wave;              // 2 x 9 arrays (p0, p1)
  DevBuf med;               // 8 arrays
  DevBuf elastic;           // rho, lambda, mu (ela_set_elastic)
  bool have_elastic = false;  // the steppers form the medium from `elastic` (F_MED3)
  double tdt = 0.0;         // 2 dtw of ela_set_elastic
  DevBuf taper;             // min_pad values
  DevBuf scratch_tm;        // time-major trace buffer
  void* scratch_dev = nullptr;  // device address of ela_scratch on cfg.device (resolved in ela_create)
  DevBuf wave2;             // Born scattered / adjoint wavefield, 2 x 9 arrays
  DevBuf coarse;            // Born coarse-time vx, vy, vz slices [slot][3][stride]
  DevBuf ckpt;              // Born adjoint checkpoints [segment][21][stride]
  DevBuf ss;                // Born secondary sources 2 x 9 arrays
  DevBuf dmod;              // Born perturbations / images, 8 arrays
  DevBuf dmod64;            // FP32 handles: the Born adjoint's FP64 image accumulator, 8 arrays
  DevBuf pos;               // converted positions (int32) and an error flag
  DevBuf staging;           // device copies of host-memory arguments
  bool have_medium = false;
  std::vector<double> filt; // 2*(sub+1)
  int64_t last_steps = 0;
  int by_x = 16;            // x extent of a tile
  int by_x_t = 0;           // x extent of the premultiplied transposed steps' tiles (0: by_x)
  int ychunk = 8;           // y planes per block
  int remap = 0;            // tile order (ela_set_tile_order)
  int xgroup = 1;           // x columns per group of the y-major order
  int64_t born_segment = 0; // Born adjoint storage: 0 auto, -1 all samples, K > 0 checkpoints every K
  int64_t born_budget = 0;  // device-memory budget of the automatic plan [bytes], 0: free memory
  int64_t born_last_segment = 0;  // plan of the last Born adjoint (-1 full storage)
  int born_physics = 0;     // Born yz term: 0 the reference's xz pattern, 1 the yz pattern (ela_set_born_mode)
  // adjoint-state gradient
  DevBuf ck1;               // outer checkpoints [n1][18][interior] (bfloat16, block-scaled int16 or native)
  DevBuf ck2;               // inner checkpoints of one outer segment [K1/K][18][interior] (native)
  DevBuf ckscale;           // per-checkpoint array scales [n1][18] + amax scratch [18]
  DevBuf drv;               // derivative combinations of one inner segment [K][9][stride]
  DevBuf g32;               // per-segment gradient partial sums [8][stride]
  DevBuf g64;               // FP64 gradient [8][stride]
  DevBuf stage2;            // host-call staging of the source-trace gradient
  int64_t grad_interval = 20;   // inner interval K (fine steps)
  int64_t grad_outer = 0;       // outer interval K1 (0: automatic)
  int grad_fmt = ELA_CK_S16;    // storage of the outer checkpoints (int16 block-scaled: inside 1e-3 of FP64)
  int64_t grad_budget = 0;      // device-memory budget of the automatic plan (0: free memory)
  int drv_fmt = ELA_CK_NATIVE;  // stored derivative combinations: ELA_CK_NATIVE or ELA_CK_S16
  bool grad_ready = false;      // a gradient forward has stored its checkpoints
  int64_t plan_K = 0, plan_K1 = 0, plan_n1 = 0;
  int plan_fmt = 0;
  // windows stacked along y (several shots in one grid)
  std::vector<int64_t> win_y0, win_ny, win_nx;
  DevBuf ytab;              // per plane {window y0, ny, nx, 0} (int32)
  DevBuf chunks;            // per y chunk {y0, y1, window y0, ny, nx, 0, 0, 0} (int32)
  int chunks_yc = 0;        // y chunk the table was built for
  int nchunks = 0;
  // graph replays of the time loops (ela_set_graph): the coarse step of the
  // nonlinear propagations, the fine-step pairs and checkpoint segments of the
  // gradient; each executable graph is kept until the launches it holds change
  // or the handle is destroyed
  bool use_graph = false;
  DevBuf counter;           // device step counters (coarse sample, fine step)
  DevBuf dfilt;             // the time filter in the handle's precision (replayed point kernels read it)
  hipStream_t cap_stream = nullptr;
  GraphCache graphs[G_COUNT];
  std::vector<unsigned char>* rec = nullptr;  // launch recorder of graph_for's signature pass
};

namespace {

template <typename T>
T* arr(DevBuf& b, int64_t stride, int k) {
  return reinterpret_cast<T*>(b.p) + stride * k;
}

template <typename V>
void sig_put(std::vector<unsigned char>& s, const V& v) {
  const unsigned char* p = reinterpret_cast<const unsigned char*>(&v);
  s.insert(s.end(), p, p + sizeof(V));
}

// Every kernel launch of the library goes through here. While the handle's
// launch recorder is set (the signature pass of a graph, see graph_for) the
// kernel, its launch shape and its argument bytes are appended to the
// recorder and nothing is launched.
template <typename... KP, typename... A>
void launch(ela_handle* h, void (*kern)(KP...), dim3 grid, dim3 block, const A&... args) {
  static_assert(sizeof...(KP) == sizeof...(A), "kernel argument count");
  std::tuple<std::decay_t<KP>...> v(static_cast<std::decay_t<KP>>(args)...);
  if (h->rec) {
    sig_put(*h->rec, reinterpret_cast<const void*>(kern));
    sig_put(*h->rec, grid);
    sig_put(*h->rec, block);
    std::apply([&](const auto&... x) { (sig_put(*h->rec, x), ...); }, v);
    return;
  }
  std::apply([&](auto&... x) {
    void* ptrs[] = {static_cast<void*>(&x)...};
    (void)hipLaunchKernel(reinterpret_cast<const void*>(kern), grid, block, ptrs, 0, h->stream);
  }, v);
}

// Destroys a cached executable graph once its launches have completed.
int drop_graph(GraphCache& gc) {
  if (!gc.exec) return ELA_OK;
  HIPCHK(hipEventSynchronize(gc.done));
  HIPCHK(hipGraphExecDestroy(gc.exec));
  gc.exec = nullptr;
  gc.sig.clear();
  return ELA_OK;
}

// Executable graph of the launches `body` issues on h->stream. The body is
// first run with the launch recorder set; if the recorded launches equal those
// the cached graph was captured from, the cached graph is returned as it is
// (the same kernels, launch shapes and arguments, hence the same work);
// otherwise the body is captured and instantiated once more and replaces the
// cache. The body must issue nothing but launches through launch() and must
// leave the host state it reads as it found it (both passes see the same
// state). The caller records gc.done on h->stream after its last replay.
template <typename F>
int graph_for(ela_handle* h, GraphCache& gc, F&& body, hipGraphExec_t* out) {
  std::vector<unsigned char> sig;
  sig.reserve(gc.sig.size());
  h->rec = &sig;
  const int rs = body();
  h->rec = nullptr;
  CHK(rs);
  if (gc.exec && sig == gc.sig) {
    *out = gc.exec;
    return ELA_OK;
  }
  CHK(drop_graph(gc));
  if (!gc.done) HIPCHK(hipEventCreateWithFlags(&gc.done, hipEventDisableTiming));
  if (!h->cap_stream) HIPCHK(hipStreamCreateWithFlags(&h->cap_stream, hipStreamNonBlocking));
  hipGraph_t g = nullptr;
  HIPCHK(hipStreamBeginCapture(h->cap_stream, hipStreamCaptureModeThreadLocal));
  hipStream_t keep = h->stream;
  h->stream = h->cap_stream;
  const int rc = body();
  h->stream = keep;
  const hipError_t ec = hipStreamEndCapture(h->cap_stream, &g);
  if (rc != ELA_OK) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (ec != hipSuccess) return fail(ELA_ERR_HIP, "graph capture failed: %s", hipGetErrorString(ec));
  hipGraphExec_t exec = nullptr;
  const hipError_t ei = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (ei != hipSuccess) return fail(ELA_ERR_HIP, "graph instantiate failed: %s", hipGetErrorString(ei));
  gc.exec = exec;
  gc.sig.swap(sig);
  *out = exec;
  return ELA_OK;
}

// Replays `exec` n times on h->stream and marks the cache's completion event.
int replay(ela_handle* h, GraphCache& gc, hipGraphExec_t exec, int64_t n) {
  for (int64_t k = 0; k < n; ++k) {
    const hipError_t e = hipGraphLaunch(exec, h->stream);
    if (e != hipSuccess) return fail(ELA_ERR_HIP, "graph launch failed: %s", hipGetErrorString(e));
  }
  HIPCHK(hipEventRecord(gc.done, h->stream));
  return ELA_OK;
}

int check_cfg(const ela_config* c) {
  if (!c) return fail(ELA_ERR_ARG, "null config");
  if (c->nz < 2 * ela::FAT + 1 || c->nx < 2 * ela::FAT + 1 || c->ny < 2 * ela::FAT + 1)
    return fail(ELA_ERR_ARG, "grid %lld x %lld x %lld too small (need > 2*FAT per axis)", (long long)c->nz,
                (long long)c->nx, (long long)c->ny);
  if (c->nts < 2) return fail(ELA_ERR_ARG, "nts must be >= 2");
  if (c->sub < 1) return fail(ELA_ERR_ARG, "sub must be >= 1");
  if (c->precision != ELA_FP32 && c->precision != ELA_FP64)
    return fail(ELA_ERR_ARG, "precision must be ELA_FP32 or ELA_FP64");
  if (!(c->dz > 0 && c->dx > 0 && c->dy > 0 && c->dts > 0)) return fail(ELA_ERR_ARG, "non-positive spacing");
  if (c->min_pad < 0) return fail(ELA_ERR_ARG, "negative min_pad");
  if (c->min_pad > std::min(c->nz, c->nx) / 2) return fail(ELA_ERR_ARG, "min_pad larger than half the grid");
  return ELA_OK;
}

int check_points(const ela_points* p, const char* what) {
  if (!p) return fail(ELA_ERR_ARG, "null %s point list", what);
  for (int g = 0; g < 7; ++g) {
    if (p->n[g] < 0) return fail(ELA_ERR_ARG, "%s grid %d: negative count", what, g);
    if (p->n[g] > (int64_t)1 << 31) return fail(ELA_ERR_ARG, "%s grid %d: too many points", what, g);
    if (p->n[g] > 0 && !p->pos[g]) return fail(ELA_ERR_ARG, "%s grid %d: null positions", what, g);
  }
  return ELA_OK;
}

template <typename T>
void set_coefs(const ela_handle* h, T* cx, T* cy, T* cz) {
  for (int k = 0; k < 4; ++k) {
    cz[k] = T(kC[k] / h->cfg.dz);
    cx[k] = T(kC[k] / h->cfg.dx);
    cy[k] = T(kC[k] / h->cfg.dx);  // reference: y coefficients use dx
  }
}

struct Tiles {
  int nbz, nbx, nby, nblocks;
};

// Chunk table of the stacked windows for the current y chunk (rebuilt when
// the y chunk changes; a synchronous upload, done before any launch of the
// propagation that uses it).
int ensure_chunks(ela_handle* h) {
  if (h->win_y0.empty() || h->chunks_yc == h->ychunk) return ELA_OK;
  std::vector<int> tab;
  for (size_t w = 0; w < h->win_y0.size(); ++w) {
    const int64_t a = h->win_y0[w] + ela::FAT, b = h->win_y0[w] + h->win_ny[w] - ela::FAT;
    for (int64_t y = a; y < b; y += h->ychunk) {
      const int row[8] = {(int)y, (int)std::min<int64_t>(y + h->ychunk, b), (int)h->win_y0[w], (int)h->win_ny[w],
                          (int)h->win_nx[w], 0, 0, 0};
      tab.insert(tab.end(), row, row + 8);
    }
  }
  CHK(h->chunks.alloc(tab.size() * sizeof(int)));
  HIPCHK(hipMemcpy(h->chunks.p, tab.data(), tab.size() * sizeof(int), hipMemcpyHostToDevice));
  h->nchunks = (int)(tab.size() / 8);
  h->chunks_yc = h->ychunk;
  return ELA_OK;
}

Tiles tiles_for(const ela_handle* h, int bx, int ychunk) {
  Tiles t;
  const int64_t iz = h->cfg.nz - 2 * ela::FAT, ix = h->cfg.nx - 2 * ela::FAT, iy = h->cfg.ny - 2 * ela::FAT;
  t.nbz = (int)((iz + ela::TZ - 1) / ela::TZ);
  t.nbx = (int)((ix + bx - 1) / bx);
  t.nby = h->win_y0.empty() ? (int)((iy + ychunk - 1) / ychunk) : h->nchunks;
  t.nblocks = t.nbz * t.nbx * t.nby;
  return t;
}

template <typename T>
ela::StepArgs<T> base_args(ela_handle* h, T* o[9], T* c[9]) {
  ela::StepArgs<T> a;
  std::memset(&a, 0, sizeof(a));
  for (int k = 0; k < 9; ++k) { a.o[k] = o[k]; a.c[k] = c[k]; }
  for (int k = 0; k < 8; ++k) a.m[k] = arr<T>(h->med, h->stride, k);
  if (h->have_elastic) {
    for (int k = 0; k < 3; ++k) a.ep[k] = arr<T>(h->elastic, h->stride, k);
    a.tdt = (T)h->tdt;
    a.tdt64 = h->tdt;
  }
  set_coefs<T>(h, a.cx, a.cy, a.cz);
  a.taper = reinterpret_cast<const T*>(h->taper.p);
  a.min_pad = (int)h->cfg.min_pad;
  a.nz = (int)h->cfg.nz; a.nx = (int)h->cfg.nx; a.ny = (int)h->cfg.ny;
  a.nzi = (int)h->nzi; a.zoff = (int)h->zoff;
  a.ys = h->cfg.nx * h->nzi;
  a.ychunk = h->ychunk;
  a.remap = h->remap;
  a.xgroup = h->xgroup;
  if (!h->win_y0.empty()) {
    (void)ensure_chunks(h);
    a.chunks = (const int*)h->chunks.p;
  }
  return a;
}

// which: bit 0 = velocity kernel, bit 1 = stress kernel; FV / FS: the
// ela::F_* work fused into the velocity / stress launch.
template <typename T, int BX, int MODE, int FV, int FS>
int launch_step_bx(ela_handle* h, ela::StepArgs<T> a, int which) {
  Tiles t = tiles_for(h, BX, a.ychunk);
  a.nbz = t.nbz; a.nbx = t.nbx; a.nby = t.nby;
  dim3 block(ela::TZ, BX);
  if (which & 1) launch(h, ela::k_velocity<T, BX, MODE, FV>, dim3(t.nblocks), block, a);
  if (which & 2) launch(h, ela::k_stress<T, BX, MODE, FS>, dim3(t.nblocks), block, a);
  return ELA_OK;
}

template <typename T, int MODE, int FV, int FS>
int launch_step(ela_handle* h, const ela::StepArgs<T>& a, int which = 3) {
  // the premultiplied transposed steps (Born adjoint, gradient) may take a tile of their own
  const int bx = MODE == ela::MODE_BORN_ADJ_PM && h->by_x_t ? h->by_x_t : h->by_x;
  switch (bx) {
    case 4: return launch_step_bx<T, 4, MODE, FV, FS>(h, a, which);
    case 8: return launch_step_bx<T, 8, MODE, FV, FS>(h, a, which);
    default: return launch_step_bx<T, 16, MODE, FV, FS>(h, a, which);
  }
}

// Forward, nonlinear-adjoint and premultiplied transposed steps form the
// medium from (rho, lambda, mu) when the handle holds them (ela_set_elastic):
// 3 medium reads per cell instead of 8. MODE_BORN_ADJ and the injections read
// the 8 scaled arrays, which k_scale_medium derives with the steppers'
// operations, so both routes give bit-identical results.
template <typename T, int MODE, int FV, int FS>
int launch_step_m(ela_handle* h, const ela::StepArgs<T>& a, int which = 3) {
  if constexpr (MODE == ela::MODE_FWD || MODE == ela::MODE_ADJ || MODE == ela::MODE_BORN_ADJ_PM) {
    if (h->have_elastic) return launch_step<T, MODE, FV | ela::F_MED3, FS | ela::F_MED3>(h, a, which);
  }
  return launch_step<T, MODE, FV, FS>(h, a, which);
}

template <typename T>
int fill(ela_handle* h, void* p, int64_t n) {
  if (n <= 0) return ELA_OK;
  const int64_t blocks = std::min<int64_t>((n + 255) / 256, 65536);
  launch(h, ela::k_fill<T>, dim3((unsigned)blocks), dim3(256), (T*)p, n, T(0));
  return ELA_OK;
}

// 3-D array [k][ny][nx][nz] (caller) <-> library layout, k = 0..n_arrays-1,
// elements of es bytes.
int copy_in_3d(ela_handle* h, void* dst, const void* src, int n_arrays, int mem, size_t es) {
  if (mem == ELA_MEM_DEVICE) CHK(check_dev_ptr(src, "3-D input array"));
  const hipMemcpyKind kind = mem == ELA_MEM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
  for (int k = 0; k < n_arrays; ++k)
    HIPCHK(hipMemcpy2DAsync((char*)dst + (k * h->stride + h->zoff) * es, h->nzi * es,
                            (const char*)src + k * h->n_model * es, h->cfg.nz * es, h->cfg.nz * es,
                            h->cfg.nx * h->cfg.ny, kind, h->stream));
  return ELA_OK;
}

int copy_in_3d(ela_handle* h, void* dst, const void* src, int n_arrays, int mem) {
  return copy_in_3d(h, dst, src, n_arrays, mem, (size_t)h->esize);
}

int copy_out_3d(ela_handle* h, void* dst, const void* src, int n_arrays, int mem, size_t es) {
  if (mem == ELA_MEM_DEVICE) CHK(check_dev_ptr(dst, "3-D output array"));
  const hipMemcpyKind kind = mem == ELA_MEM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
  for (int k = 0; k < n_arrays; ++k)
    HIPCHK(hipMemcpy2DAsync((char*)dst + k * h->n_model * es, h->cfg.nz * es,
                            (const char*)src + (k * h->stride + h->zoff) * es, h->nzi * es, h->cfg.nz * es,
                            h->cfg.nx * h->cfg.ny, kind, h->stream));
  return ELA_OK;
}

int copy_out_3d(ela_handle* h, void* dst, const void* src, int n_arrays, int mem) {
  return copy_out_3d(h, dst, src, n_arrays, mem, (size_t)h->esize);
}

// ---- point lists ------------------------------------------------------------

struct DevPoints {
  const int32_t* pos[7];
  int64_t n[7];
};

// Copies (host) or reads (device) the caller's int64 positions, converts them
// to internal int32 offsets on the GPU and checks that every point lies in the
// FD interior.
// Bytes of scratch that convert_points uses for `total` points of a
// device-memory call (int32 positions and the flag).
size_t point_scratch_bytes(int64_t total) { return (size_t)round_up(total * 4, 256) + 256; }

int convert_points(ela_handle* h, const ela_points* sp, const ela_points* rp, int mem, DevPoints& ds, DevPoints& dr,
                   void* scratch = nullptr) {
  int64_t total = 0;
  for (int g = 0; g < 7; ++g) total += sp->n[g] + rp->n[g];
  if (mem == ELA_MEM_DEVICE)
    for (int g = 0; g < 7; ++g) {
      if (sp->n[g]) CHK(check_dev_ptr(sp->pos[g], "a source position list"));
      if (rp->n[g]) CHK(check_dev_ptr(rp->pos[g], "a receiver position list"));
    }
  // layout: [int64 staging (host calls)] [int32 converted] [int flag]; with
  // `scratch` (device calls only) the same without the staging, in the caller's scratch
  if (scratch && mem != ELA_MEM_DEVICE) return fail(ELA_ERR_ARG, "scratch point conversion needs device positions");
  const size_t b64 = scratch ? 0 : round_up(total * 8, 256), b32 = round_up(total * 4, 256);
  if (!scratch) CHK(h->pos.alloc(b64 + b32 + 256));
  char* base = scratch ? (char*)scratch : (char*)h->pos.p;
  int64_t* s64 = (int64_t*)base;
  int32_t* s32 = (int32_t*)(base + b64);
  int* flag = (int*)(base + b64 + b32);
  launch(h, ela::k_fill<int>, dim3(1), dim3(64), flag, (int64_t)1, 0);
  int64_t off = 0;
  for (int side = 0; side < 2; ++side) {
    const ela_points* p = side ? rp : sp;
    DevPoints& d = side ? dr : ds;
    for (int g = 0; g < 7; ++g) {
      const int64_t n = p->n[g];
      d.n[g] = n;
      d.pos[g] = s32 + off;
      if (n) {
        const int64_t* src = p->pos[g];
        if (mem == ELA_MEM_HOST) {
          HIPCHK(hipMemcpyAsync(s64 + off, p->pos[g], n * 8, hipMemcpyHostToDevice, h->stream));
          src = s64 + off;
        }
        launch(h, ela::k_convert_pos, dim3((unsigned)((n + 255) / 256)), dim3(256), src,
                           s32 + off, n, (int)h->cfg.nz, (int)h->cfg.nx, (int)h->cfg.ny, (int)h->nzi, (int)h->zoff,
                           h->win_y0.empty() ? nullptr : (const int*)h->ytab.p, flag);
      }
      off += n;
    }
  }
  int bad = 0;
  HIPCHK(hipMemcpyAsync(&bad, flag, sizeof(int), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  if (bad) return fail(ELA_ERR_ARG, "a source or receiver position lies outside the FD interior");
  return ELA_OK;
}

template <typename T>
ela::PointArgs<T> make_points(ela_handle* h, const DevPoints& pts, T* const field[9], T* const tr[9],
                              bool time_major, int64_t tm_cols) {
  ela::PointArgs<T> a;
  std::memset(&a, 0, sizeof(a));
  int64_t off = 0;
  for (int c = 0; c < 9; ++c) {
    const int g = kCompGrid[c];
    a.pos[c] = pts.pos[g];
    a.n[c] = pts.n[g];
    a.off[c] = off;
    a.field[c] = field[c];
    a.tr[c] = time_major ? tr[0] + off : tr[c];
    off += pts.n[g];
  }
  a.off[9] = off;
  a.tr_stride_pt = time_major ? 1 : h->cfg.nts;
  a.tr_stride_t = time_major ? tm_cols : 1;
  a.taper = reinterpret_cast<const T*>(h->taper.p);
  a.min_pad = (int)h->cfg.min_pad;
  a.nz = (int)h->cfg.nz; a.nx = (int)h->cfg.nx; a.ny = (int)h->cfg.ny;
  a.nzi = (int)h->nzi; a.zoff = (int)h->zoff;
  a.ytab = h->win_y0.empty() ? nullptr : (const int*)h->ytab.p;
  return a;
}

// Host time arguments of a point launch. A launch that reads its coarse sample
// (its_dev) or its fine step (step_dev) from device memory keeps them out of
// its arguments, so that the replayed graph's launches do not depend on them.
template <typename T>
void set_time(ela::PointArgs<T>& a, int its, double fa, double fb) {
  a.its = a.its_dev || a.step_dev ? 0 : its;
  a.fa = a.step_dev ? T(0) : T(fa);
  a.fb = a.step_dev ? T(0) : T(fb);
}

// The point launch reads its fine step (*kdev + off) from device memory and
// its weights from the handle's filter table (offset fo: 0 for filt[i2], 1
// for filt[i2 + 1]).
template <typename T>
void set_step(ela_handle* h, ela::PointArgs<T>& a, const int* kdev, int off, int fo) {
  a.step_dev = kdev;
  a.step_off = off;
  a.filt = static_cast<const T*>(h->dfilt.p);
  a.sub = (int)h->cfg.sub;
  a.fo = fo;
}

// A replayed step's position relative to the device step counter, and
// whether the step advances the counter afterwards (the last step of a
// replayed pair).
struct DevStep {
  int* k = nullptr;   // device counter, or null (host index)
  int off = 0;
  int advance = 0;
};

template <typename T, bool ACC, bool GRAD = false, bool PM = false>
void launch_inject(ela_handle* h, ela::PointArgs<T> a, int its, double fa, double fb) {
  if (a.off[9] == 0) return;
  set_time(a, its, fa, fb);
  const unsigned nb = (unsigned)((a.off[9] + 255) / 256);
  launch(h, ela::k_inject<T, ACC, GRAD, PM>, dim3(nb), dim3(256), a);
}

template <typename T>
void launch_record(ela_handle* h, ela::PointArgs<T> a, int its, double fa, double fb) {
  if (a.off[9] == 0) return;
  set_time(a, its, fa, fb);
  const unsigned nb = (unsigned)((a.off[9] + 255) / 256);
  launch(h, ela::k_record<T>, dim3(nb), dim3(256), a);
}

template <typename T>
void launch_transpose(ela_handle* h, const T* in, T* out, int64_t rows, int64_t cols, int64_t ld_in, int64_t ld_out) {
  if (rows <= 0 || cols <= 0) return;
  dim3 grid((unsigned)((cols + 31) / 32), (unsigned)((rows + 31) / 32));
  launch(h, ela::k_transpose<T>, grid, dim3(256), in, out, rows, cols, ld_in, ld_out);
}

// ---- trace staging for host-memory calls ------------------------------------

int64_t trace_elems(const ela_points* p, int64_t nts) {
  int64_t n = 0;
  for (int c = 0; c < 9; ++c) n += p->n[kCompGrid[c]] * nts;
  return n;
}

// Device pointers for the 9 source and 9 receiver trace arrays: the caller's
// (device calls) or library staging copies (host calls).
int stage_traces(ela_handle* h, const ela_points* sp, const ela_traces* st, const ela_points* rp,
                 const ela_traces* rt, bool copy_src, bool copy_rec, int mem, void* dsrc[9], void* drec[9]) {
  if (mem == ELA_MEM_DEVICE) {
    for (int c = 0; c < 9; ++c) {
      if (sp->n[kCompGrid[c]]) CHK(check_dev_ptr(st->tr[c], "a source trace array"));
      if (rp->n[kCompGrid[c]]) CHK(check_dev_ptr(rt->tr[c], "a receiver trace array"));
      dsrc[c] = st->tr[c];
      drec[c] = rt->tr[c];
    }
    return ELA_OK;
  }
  const int64_t nts = h->cfg.nts;
  const size_t bytes = (trace_elems(sp, nts) + trace_elems(rp, nts)) * h->esize + 18 * 256;
  CHK(h->staging.alloc(bytes));
  char* base = (char*)h->staging.p;
  size_t off = 0;
  for (int c = 0; c < 9; ++c) {
    const int64_t ns = sp->n[kCompGrid[c]] * nts, nr = rp->n[kCompGrid[c]] * nts;
    dsrc[c] = base + off;
    off += round_up(ns * h->esize, 256);
    drec[c] = base + off;
    off += round_up(nr * h->esize, 256);
    if (copy_src && ns) HIPCHK(hipMemcpyAsync(dsrc[c], st->tr[c], ns * h->esize, hipMemcpyHostToDevice, h->stream));
    if (copy_rec && nr) HIPCHK(hipMemcpyAsync(drec[c], rt->tr[c], nr * h->esize, hipMemcpyHostToDevice, h->stream));
  }
  return ELA_OK;
}

int stage_out_traces(ela_handle* h, const ela_points* p, void* const dev[9], const ela_traces* host) {
  for (int c = 0; c < 9; ++c) {
    const int64_t n = p->n[kCompGrid[c]] * h->cfg.nts;
    if (n) HIPCHK(hipMemcpyAsync(host->tr[c], dev[c], n * h->esize, hipMemcpyDeviceToHost, h->stream));
  }
  return ELA_OK;
}

// ---- the fine-step loops --------------------------------------------------

// Two time levels (p0 = level n-1, overwritten by n+1; p1 = level n) and,
// for the premultiplied Born-adjoint stepper, their medium products q0, q1.
template <typename T>
struct Wave {
  T* p0[9];
  T* p1[9];
  T* q0[9];
  T* q1[9];
  void swap() {
    for (int k = 0; k < 9; ++k) {
      std::swap(p0[k], p1[k]);
      std::swap(q0[k], q1[k]);
    }
  }
};

template <typename T>
Wave<T> wave_ptrs(DevBuf& b, int64_t stride, bool premultiplied = false) {
  Wave<T> w;
  for (int k = 0; k < 9; ++k) {
    w.p0[k] = arr<T>(b, stride, k);
    w.p1[k] = arr<T>(b, stride, 9 + k);
    w.q0[k] = premultiplied ? arr<T>(b, stride, 18 + k) : nullptr;
    w.q1[k] = premultiplied ? arr<T>(b, stride, 27 + k) : nullptr;
  }
  return w;
}

// Premultiplied Born-adjoint step arguments: read M c from q1, write M n to q0.
template <typename T>
void set_pm(ela_handle* h, ela::StepArgs<T>& a, const Wave<T>& w) {
  for (int k = 0; k < 9; ++k) { a.q[k] = w.q1[k]; a.qn[k] = w.q0[k]; }
  (void)h;
}

template <typename T>
void set_pm(ela_handle* h, ela::PointArgs<T>& a, const Wave<T>& w) {
  for (int k = 0; k < 9; ++k) a.q[k] = w.q0[k];
  for (int k = 0; k < 8; ++k) a.m[k] = arr<T>(h->med, h->stride, k);
}

// The 9 components of a time-major trace buffer share one base pointer
// (make_points offsets each component).
template <typename T>
struct TmViews {
  T* v[9];
  explicit TmViews(T* tm) {
    for (int c = 0; c < 9; ++c) v[c] = tm;
  }
};

// Interior cells the stepper advances (the windows' interiors when windows
// are stacked).
int64_t interior_cells(const ela_handle* h) {
  if (h->win_y0.empty())
    return (h->cfg.nz - 2 * ela::FAT) * (h->cfg.nx - 2 * ela::FAT) * (h->cfg.ny - 2 * ela::FAT);
  int64_t n = 0;
  for (size_t w = 0; w < h->win_y0.size(); ++w)
    n += (h->cfg.nz - 2 * ela::FAT) * (h->win_nx[w] - 2 * ela::FAT) * (h->win_ny[w] - 2 * ela::FAT);
  return n;
}

int64_t total_points(const DevPoints& p) {
  int64_t n = 0;
  for (int c = 0; c < 9; ++c) n += p.n[kCompGrid[c]];
  return n;
}

// `scratch` (device-memory forward calls only): the device address of
// ela_scratch; the point lists, flag and time-major receiver buffer are
// placed there and nothing is allocated (eager launches, no graph).
template <typename T>
int nl_run(ela_handle* h, const ela_points* sp_in, const ela_traces* st_in, const ela_points* rp_in,
           const ela_traces* rt_in, bool adjoint, int mem, void* scratch = nullptr) {
  if (!h->have_medium) return fail(ELA_ERR_STATE, "medium not set");
  CHK(check_points(sp_in, "source"));
  CHK(check_points(rp_in, "receiver"));
  const int64_t nts = h->cfg.nts, sub = h->cfg.sub, hh = sub + 1;
  int64_t npts = 0;
  for (int g = 0; g < 7; ++g) npts += sp_in->n[g] + rp_in->n[g];
  size_t tm_off = 0;
  if (scratch) {
    if (adjoint || mem != ELA_MEM_DEVICE) return fail(ELA_ERR_ARG, "the scratch path is the device-memory forward");
    if (!h->win_y0.empty() && h->chunks_yc != h->ychunk)
      return fail(ELA_ERR_STATE, "stacked windows: call ela_set_windows again after ela_set_tiling");
    int64_t nrec = 0;
    for (int c = 0; c < 9; ++c) nrec += rp_in->n[kCompGrid[c]];
    tm_off = point_scratch_bytes(npts);
    const size_t need = tm_off + (size_t)std::max<int64_t>(1, nrec * nts) * sizeof(T);
    if (need > kScratchBytes)
      return fail(ELA_ERR_ARG, "ela_nl_forward_noalloc: %zu bytes of scratch needed (%lld points, %lld receiver "
                  "values x %lld samples), the limit is %zu", need, (long long)npts, (long long)nrec,
                  (long long)nts, kScratchBytes);
  }
  DevPoints sp, rp;
  CHK(convert_points(h, sp_in, rp_in, mem, sp, rp, scratch));
  void* ds[9];
  void* dr[9];
  CHK(stage_traces(h, sp_in, st_in, rp_in, rt_in, !adjoint, adjoint, mem, ds, dr));
  T* src_tr[9];
  T* rec_tr[9];
  for (int c = 0; c < 9; ++c) { src_tr[c] = (T*)ds[c]; rec_tr[c] = (T*)dr[c]; }

  // time-major receiver buffer [nts][nrec_total]: coalesced recording
  const int64_t nrec_tot = total_points(rp);
  if (!scratch) CHK(h->scratch_tm.alloc(std::max<int64_t>(1, nrec_tot * nts) * sizeof(T)));
  T* tm = scratch ? (T*)((char*)scratch + tm_off) : (T*)h->scratch_tm.p;
  T* tmv[9];
  for (int c = 0; c < 9; ++c) tmv[c] = tm;

  Wave<T> w = wave_ptrs<T>(h->wave, h->stride);
  CHK(fill<T>(h, h->wave.p, 18 * h->stride));
  if (!adjoint) {
    CHK(fill<T>(h, tm, nrec_tot * nts));
  } else {
    int64_t off = 0;
    for (int c = 0; c < 9; ++c) {
      const int64_t n = rp.n[kCompGrid[c]];
      launch_transpose<T>(h, rec_tr[c], tm + off, n, nts, nts, nrec_tot);
      off += n;
    }
    for (int c = 0; c < 9; ++c) CHK(fill<T>(h, src_tr[c], sp.n[kCompGrid[c]] * nts));
  }

  const double* f = h->filt.data();
  // One coarse step: sub fine steps (stepper, injection, recording). With
  // its_dev the coarse index is read from device memory (graph replays).
  auto coarse_step = [&](int64_t its, const int* its_dev) -> int {
    if (!adjoint) {
      for (int64_t it2 = 1; it2 <= sub; ++it2) {
        ela::StepArgs<T> a = base_args<T>(h, w.p0, w.p1);
        CHK((launch_step_m<T, ela::MODE_FWD, 0, 0>(h, a)));
        ela::PointArgs<T> S = make_points<T>(h, sp, w.p0, src_tr, false, 0);
        ela::PointArgs<T> R = make_points<T>(h, rp, w.p0, tmv, true, nrec_tot);
        S.its_dev = R.its_dev = its_dev;
        launch_inject<T, false>(h, S, (int)its, f[it2 - 1], f[hh + it2 - 1]);
        launch_record<T>(h, R, (int)its, f[it2], f[hh + it2]);
        w.swap();
      }
    } else {
      for (int64_t it2 = sub - 1; it2 >= 0; --it2) {
        ela::StepArgs<T> a = base_args<T>(h, w.p0, w.p1);
        CHK((launch_step_m<T, ela::MODE_ADJ, 0, 0>(h, a)));
        ela::PointArgs<T> R = make_points<T>(h, rp, w.p0, tmv, true, nrec_tot);
        ela::PointArgs<T> S = make_points<T>(h, sp, w.p0, src_tr, false, 0);
        S.its_dev = R.its_dev = its_dev;
        launch_inject<T, false>(h, R, (int)its, f[it2 + 1], f[hh + it2 + 1]);
        launch_record<T>(h, S, (int)its, f[it2], f[hh + it2]);
        w.swap();
      }
    }
    return ELA_OK;
  };
  // sub is even (the doubled sub-sampling) whenever the graph is used, so the
  // wavefield pointers are back in place after one coarse step and the same
  // captured step serves every coarse sample
  const bool graph = h->use_graph && sub % 2 == 0 && !scratch;
  hipGraphExec_t exec = nullptr;
  if (graph) {
    CHK(ensure_chunks(h));  // no synchronous upload inside the capture
    CHK(h->counter.alloc(256));
    int* cnt = (int*)h->counter.p;
    auto body = [&]() -> int {
      CHK(coarse_step(0, cnt));
      launch(h, ela::k_count, dim3(1), dim3(64), cnt, adjoint ? -1 : 1);
      return ELA_OK;
    };
    CHK(graph_for(h, h->graphs[adjoint ? G_NL_ADJ : G_NL], body, &exec));
    launch(h, ela::k_fill<int>, dim3(1), dim3(64), cnt, (int64_t)1, adjoint ? (int)(nts - 2) : 0);
  }
  HIPCHK(hipEventRecord(h->ev0, h->stream));
  if (graph) {
    CHK(replay(h, h->graphs[adjoint ? G_NL_ADJ : G_NL], exec, nts - 1));
  } else if (!adjoint) {
    for (int64_t its = 0; its < nts - 1; ++its) CHK(coarse_step(its, nullptr));
  } else {
    for (int64_t its = nts - 2; its >= 0; --its) CHK(coarse_step(its, nullptr));
  }
  HIPCHK(hipEventRecord(h->ev1, h->stream));
  HIPCHK(hipGetLastError());
  h->last_steps = (nts - 1) * sub;

  if (!adjoint) {
    int64_t off = 0;
    for (int c = 0; c < 9; ++c) {
      const int64_t n = rp.n[kCompGrid[c]];
      launch_transpose<T>(h, tm + off, rec_tr[c], nts, n, nrec_tot, nts);
      off += n;
    }
  }
  HIPCHK(hipGetLastError());
  if (mem == ELA_MEM_HOST) {
    if (!adjoint) CHK(stage_out_traces(h, rp_in, dr, rt_in));
    else CHK(stage_out_traces(h, sp_in, ds, st_in));
    HIPCHK(hipStreamSynchronize(h->stream));
  }
  return ELA_OK;
}

// ---- Born operators ----------------------------------------------------------
//
// Pass 1 propagates the background source and accumulates the coarse-time
// vx, vy, vz (interpWavefieldVxVyVz_3D, BornElasticGpuFunctions_3D.cu:1268-1316):
// fine step it2 of coarse step k adds its velocities to V[k] and V[k+1] with
// the hat weights, so V[k] is final once coarse step k has run. Pass 2 needs
// the secondary sources of coarse samples its and its+1, i.e. V[its-1..its+2].
//
// Forward: the two passes are interleaved, the background running two coarse
// samples ahead of the scattered field, so only four coarse slices are ever
// live (the reference stores all nts of them in pinned host memory; at
// 72 M cells and 501 samples that is 436 GB).
// Adjoint: the scattered field runs backwards in time and needs V in reverse
// order. V is kept for all samples when it fits in the memory budget;
// otherwise the forward sweep stores exact (FP32/FP64) checkpoints of the
// background state plus the partial V of the segment start every K coarse
// samples, and each segment is recomputed before its adjoint sweep. The
// kernels and their order are the same in all cases, so the results are
// bit-identical.

// One coarse step k of pass 1 (sub fine steps): VL = V[k], VR = V[k+1].
template <typename T>
int bg_coarse_step(ela_handle* h, Wave<T>& w, const DevPoints& sp, T* const src_tr[9], int64_t k, T* VL, T* VR) {
  const int64_t sub = h->cfg.sub, hh = sub + 1, np = h->stride;
  const double* f = h->filt.data();
  for (int64_t it2 = 1; it2 <= sub; ++it2) {
    ela::StepArgs<T> a = base_args<T>(h, w.p0, w.p1);
    for (int c = 0; c < 3; ++c) { a.accl[c] = VL + c * np; a.accr[c] = VR + c * np; }
    a.accfl = T(f[it2]); a.accfr = T(f[hh + it2]);
    CHK((launch_step_m<T, ela::MODE_FWD, ela::F_ACC, 0>(h, a)));
    ela::PointArgs<T> S = make_points<T>(h, sp, w.p0, src_tr, false, 0);
    for (int c = 0; c < 3; ++c) { S.accl[c] = a.accl[c]; S.accr[c] = a.accr[c]; }
    S.accfl = a.accfl; S.accfr = a.accfr;
    launch_inject<T, true>(h, S, (int)k, f[it2 - 1], f[hh + it2 - 1]);
    w.swap();
  }
  return ELA_OK;
}

// Secondary sources (forward, ss = out) or imaging (adjoint, dmo += ...) of
// coarse sample its; Vo, Vc, Vn are the 3-field slices V[its-1], V[its],
// V[its+1] (null outside 0..nts-1).
template <typename T>
ela::ImagingArgs<T> imaging_args(ela_handle* h, const T* Vo, const T* Vc, const T* Vn, int64_t its, const T* dm,
                                 double* dmo, T* const ss[9]) {
  ela::ImagingArgs<T> a;
  std::memset(&a, 0, sizeof(a));
  const int64_t np = h->stride;
  for (int k = 0; k < 3; ++k) {
    a.o[k] = Vo ? Vo + k * np : nullptr;
    a.c[k] = Vc + k * np;
    a.n[k] = Vn ? Vn + k * np : nullptr;
  }
  for (int k = 0; k < 8; ++k) {
    if (dm) a.dm[k] = dm + k * np;
    if (dmo) a.dmo[k] = dmo + k * np;
  }
  for (int k = 0; k < 9; ++k) a.ss[k] = ss[k];
  set_coefs<T>(h, a.cx, a.cy, a.cz);
  a.dts_inv = T(1.0 / (2.0 * h->cfg.dts));
  a.its = (int)its; a.nts = (int)h->cfg.nts;
  a.physics = h->born_physics;
  a.nz = (int)h->cfg.nz; a.nx = (int)h->cfg.nx; a.ny = (int)h->cfg.ny;
  a.nzi = (int)h->nzi; a.zoff = (int)h->zoff;
  a.ys = h->cfg.nx * h->nzi;
  a.ychunk = h->ychunk;
  a.remap = h->remap;
  a.xgroup = h->xgroup;
  if (!h->win_y0.empty()) {
    (void)ensure_chunks(h);
    a.chunks = (const int*)h->chunks.p;
  }
  Tiles t = tiles_for(h, 8, a.ychunk);
  a.nbz = t.nbz; a.nbx = t.nbx; a.nby = t.nby;
  return a;
}

template <typename T, bool ADJ>
void launch_imaging(ela_handle* h, const ela::ImagingArgs<T>& a) {
  Tiles t = tiles_for(h, 8, a.ychunk);
  launch(h, ela::k_imaging<T, 8, ADJ>, dim3(t.nblocks), dim3(ela::TZ, 8), a);
}

// Storage plan of the Born adjoint's coarse wavefield.
struct BornPlan {
  bool full = true;      // V for all nts samples
  int64_t K = 0;         // checkpoint segment length [coarse samples]
  int64_t S = 0;         // number of segments
  int64_t slots = 0;     // V slices held (full: nts; checkpointed: K + 3)
};
// A checkpointed plan cuts the nts - 1 coarse steps into S segments aligned
// to the end of the record (segment s starts at born_seg_start; all but the
// first hold K steps) and stores S - 1 checkpoints: the forward sweep leaves
// the last segment's V slices in the window, where the adjoint sweep starts.

// Arrays (of h->stride elements) the Born adjoint needs besides V and the
// checkpoints: medium 8, background wavefield 18, adjoint wavefield with its
// premultiplied levels 36, secondary sources 18, images 8 (FP32 handles: and
// their FP64 accumulator, 16 more).
int64_t born_seg_start(const BornPlan& p, int64_t ncoarse, int64_t s) {
  return std::max<int64_t>(0, ncoarse - (p.S - s) * p.K);
}

int64_t born_fixed_arrays(const ela_handle* h) { return 8 + 18 + 36 + 18 + 8 + (h->esize == 4 ? 16 : 0); }

int born_plan(ela_handle* h, BornPlan& p) {
  const int64_t nts = h->cfg.nts, ncoarse = nts - 1;
  const double arr_b = double(h->stride) * h->esize;
  if (h->born_segment < 0) {
    p.full = true; p.slots = nts; return ELA_OK;
  }
  if (h->born_segment > 0) {
    p.full = false;
    p.K = std::min<int64_t>(h->born_segment, ncoarse);
    p.S = (ncoarse + p.K - 1) / p.K;
    p.slots = p.K + 3;
    return ELA_OK;
  }
  // automatic: full storage if it fits, else the longest segment that fits
  double budget;
  if (h->born_budget > 0) {
    budget = double(h->born_budget);
  } else {
    size_t fr = 0, tot = 0;
    HIPCHK(hipMemGetInfo(&fr, &tot));
    // the gradient's step buffers count as free: the Born adjoint releases them before it allocates
    budget = double(fr) + double(h->coarse.bytes) + double(h->ckpt.bytes) + double(h->wave2.bytes) +
             double(h->ss.bytes) + double(h->dmod.bytes) + double(h->dmod64.bytes) + double(h->wave.bytes) +
             double(h->med.bytes) + double(h->drv.bytes) + double(h->g32.bytes) + double(h->g64.bytes) -
             double(2ull << 30);  // 2 GiB margin for traces and the runtime
  }
  const double avail = budget / arr_b - double(born_fixed_arrays(h));
  if (double(3 * nts) <= avail) { p.full = true; p.slots = nts; return ELA_OK; }
  for (int64_t K = ncoarse; K >= 1; --K) {
    const int64_t S = (ncoarse + K - 1) / K;
    if (double(21 * (S - 1) + 3 * (K + 3)) <= avail) {
      p.full = false; p.K = K; p.S = S; p.slots = K + 3;
      return ELA_OK;
    }
  }
  return fail(ELA_ERR_MEM, "Born adjoint: %.1f GB of device memory cannot hold the checkpoints of %lld coarse samples",
              budget / 1e9, (long long)nts);
}

template <typename T>
int born_forward_run(ela_handle* h, const DevPoints& sp, T* const src_tr[9], const DevPoints& rp, T* tm,
                     int64_t nrec_tot, const T* dm) {
  const int64_t nts = h->cfg.nts, sub = h->cfg.sub, hh = sub + 1, np = h->stride;
  CHK(h->coarse.alloc(4 * 3 * np * sizeof(T)));
  CHK(h->wave2.alloc(18 * np * sizeof(T)));
  CHK(h->ss.alloc(18 * np * sizeof(T)));
  T* ring = (T*)h->coarse.p;
  auto V = [&](int64_t t) -> T* { return (t < 0 || t >= nts) ? nullptr : ring + (t % 4) * 3 * np; };
  T* ssL[9];
  T* ssR[9];
  for (int k = 0; k < 9; ++k) { ssL[k] = arr<T>(h->ss, np, k); ssR[k] = arr<T>(h->ss, np, 9 + k); }
  Wave<T> bg = wave_ptrs<T>(h->wave, np), sc = wave_ptrs<T>(h->wave2, np);
  CHK(fill<T>(h, h->wave.p, 18 * np));
  CHK(fill<T>(h, h->wave2.p, 18 * np));
  CHK(fill<T>(h, h->ss.p, 18 * np));
  CHK(fill<T>(h, ring, 4 * 3 * np));
  CHK(fill<T>(h, tm, nrec_tot * nts));
  const double* f = h->filt.data();

  auto scattered_step = [&](int64_t its) -> int {
    if (its == 0) launch_imaging<T, false>(h, imaging_args<T>(h, V(-1), V(0), V(1), 0, dm, nullptr, ssL));
    launch_imaging<T, false>(h, imaging_args<T>(h, V(its), V(its + 1), V(its + 2), its + 1, dm, nullptr, ssR));
    for (int64_t it2 = 1; it2 <= sub; ++it2) {
      ela::StepArgs<T> a = base_args<T>(h, sc.p0, sc.p1);
      for (int k = 0; k < 9; ++k) { a.ssl[k] = ssL[k]; a.ssr[k] = ssR[k]; }
      a.ssfl = T(f[it2 - 1]); a.ssfr = T(f[hh + it2 - 1]);
      CHK((launch_step_m<T, ela::MODE_FWD, ela::F_SS, ela::F_SS>(h, a)));
      launch_record<T>(h, make_points<T>(h, rp, sc.p0, TmViews<T>(tm).v, true, nrec_tot), (int)its, f[it2],
                       f[hh + it2]);
      sc.swap();
    }
    for (int k = 0; k < 9; ++k) std::swap(ssL[k], ssR[k]);
    return ELA_OK;
  };

  int64_t next = 0;  // next scattered coarse step
  for (int64_t k = 0; k <= nts - 2; ++k) {
    if (k >= 3) CHK(fill<T>(h, V(k + 1), 3 * np));  // reuse the slice of V[k-3]
    CHK(bg_coarse_step<T>(h, bg, sp, src_tr, k, V(k), V(k + 1)));
    // V[0..k] are final (all of V once k == nts-2)
    while (next <= nts - 2 && (next + 2 <= k || k == nts - 2)) CHK(scattered_step(next++));
  }
  h->last_steps = 2 * (nts - 1) * sub;
  return ELA_OK;
}

template <typename T>
int born_adjoint_run(ela_handle* h, const DevPoints& sp, T* const src_tr[9], const DevPoints& rp, T* tm,
                     int64_t nrec_tot, double* dm) {
  const int64_t nts = h->cfg.nts, sub = h->cfg.sub, hh = sub + 1, np = h->stride;
  BornPlan plan;
  CHK(born_plan(h, plan));
  h->born_last_segment = plan.full ? -1 : plan.K;
  // release what the new plan does not reuse before allocating (peak memory); the gradient's
  // step buffers, which the plan counted as free, are allocated again by the next gradient
  const size_t ck_bytes = plan.full ? 0 : (size_t)(plan.S - 1) * 21 * np * sizeof(T);
  const size_t v_bytes = (size_t)plan.slots * 3 * np * sizeof(T);
  if (h->ckpt.bytes != ck_bytes) h->ckpt.release();
  if (h->coarse.bytes != v_bytes) h->coarse.release();
  if (h->born_segment == 0 && h->born_budget <= 0) {  // the automatic plan counted them as free
    h->drv.release();
    h->g32.release();
    h->g64.release();
  }
  CHK(h->ckpt.alloc(ck_bytes));
  CHK(h->coarse.alloc(v_bytes));
  CHK(h->wave2.alloc(36 * np * sizeof(T)));
  CHK(h->ss.alloc(18 * np * sizeof(T)));
  T* pool = (T*)h->coarse.p;
  T* ck = (T*)h->ckpt.p;
  T* ssL[9];
  T* ssR[9];
  for (int k = 0; k < 9; ++k) { ssL[k] = arr<T>(h->ss, np, k); ssR[k] = arr<T>(h->ss, np, 9 + k); }
  Wave<T> bg = wave_ptrs<T>(h->wave, np), ad = wave_ptrs<T>(h->wave2, np, true);
  const double* f = h->filt.data();
  int64_t steps = 0;

  // V slice of coarse sample t in the current storage window [base, base + slots)
  int64_t base = 0;
  auto V = [&](int64_t t) -> T* {
    return (t < 0 || t >= nts) ? nullptr : pool + (t - base) * 3 * np;
  };

  CHK(fill<T>(h, h->wave.p, 18 * np));
  if (plan.full) {
    CHK(fill<T>(h, pool, nts * 3 * np));
    for (int64_t k = 0; k <= nts - 2; ++k) CHK(bg_coarse_step<T>(h, bg, sp, src_tr, k, V(k), V(k + 1)));
    steps += (nts - 1) * sub;
  } else {
    // forward sweep: checkpoint s < S - 1 holds (p0, p1) and the partial
    // V[k0] at the start of segment s (first coarse step k0); two rotating
    // slices carry the accumulation. From the last segment's first step kl on,
    // the samples accumulate in the window of base kl instead and stay there:
    // the adjoint sweep starts with that segment, which therefore needs no
    // recomputation.
    T* rot[2] = {pool, pool + 3 * np};
    const int64_t ncoarse = nts - 1, kl = born_seg_start(plan, ncoarse, plan.S - 1);
    CHK(fill<T>(h, pool, 6 * np));
    for (int64_t s = 0; s < plan.S - 1; ++s) {
      const int64_t k0 = born_seg_start(plan, ncoarse, s), k1 = born_seg_start(plan, ncoarse, s + 1);
      T* c = ck + s * 21 * np;
      for (int f9 = 0; f9 < 9; ++f9) {
        HIPCHK(hipMemcpyAsync(c + f9 * np, bg.p0[f9], np * sizeof(T), hipMemcpyDeviceToDevice, h->stream));
        HIPCHK(hipMemcpyAsync(c + (9 + f9) * np, bg.p1[f9], np * sizeof(T), hipMemcpyDeviceToDevice, h->stream));
      }
      HIPCHK(hipMemcpyAsync(c + 18 * np, rot[k0 % 2], 3 * np * sizeof(T), hipMemcpyDeviceToDevice, h->stream));
      for (int64_t k = k0; k < k1; ++k) {
        CHK(bg_coarse_step<T>(h, bg, sp, src_tr, k, rot[k % 2], rot[(k + 1) % 2]));
        CHK(fill<T>(h, rot[k % 2], 3 * np));
      }
    }
    // the partial V[kl] (in rot[kl % 2]) to slot 0, the slots of V[kl + 1 .. nts - 1] cleared
    base = kl;
    if (rot[kl % 2] != V(kl))
      HIPCHK(hipMemcpyAsync(V(kl), rot[kl % 2], 3 * np * sizeof(T), hipMemcpyDeviceToDevice, h->stream));
    CHK(fill<T>(h, V(kl + 1), (nts - 1 - kl) * 3 * np));
    for (int64_t k = kl; k <= nts - 2; ++k) CHK(bg_coarse_step<T>(h, bg, sp, src_tr, k, V(k), V(k + 1)));
    steps += (nts - 1) * sub;
  }

  // adjoint sweep
  CHK(fill<T>(h, h->wave2.p, 36 * np));
  CHK(fill<T>(h, h->ss.p, 18 * np));
  auto adjoint_step = [&](int64_t its) -> int {
    for (int64_t it2 = sub - 1; it2 >= 0; --it2) {
      ela::StepArgs<T> a = base_args<T>(h, ad.p0, ad.p1);
      set_pm<T>(h, a, ad);
      for (int k = 0; k < 9; ++k) { a.accl[k] = ssL[k]; a.accr[k] = ssR[k]; }
      a.accfl = T(f[it2]); a.accfr = T(f[hh + it2]);
      CHK((launch_step_m<T, ela::MODE_BORN_ADJ_PM, ela::F_ACC, ela::F_ACC>(h, a)));
      ela::PointArgs<T> R = make_points<T>(h, rp, ad.p0, TmViews<T>(tm).v, true, nrec_tot);
      set_pm<T>(h, R, ad);
      for (int k = 0; k < 9; ++k) { R.accl[k] = ssL[k]; R.accr[k] = ssR[k]; }
      R.accfl = a.accfl; R.accfr = a.accfr;
      launch_inject<T, true, false, true>(h, R, (int)its, f[it2 + 1], f[hh + it2 + 1]);
      ad.swap();
    }
    launch_imaging<T, true>(h, imaging_args<T>(h, V(its), V(its + 1), V(its + 2), its + 1, nullptr, dm, ssR));
    for (int k = 0; k < 9; ++k) std::swap(ssL[k], ssR[k]);
    for (int k = 0; k < 9; ++k) CHK(fill<T>(h, ssL[k], np));
    steps += sub;
    return ELA_OK;
  };

  if (plan.full) {
    for (int64_t its = nts - 2; its >= 0; --its) CHK(adjoint_step(its));
  } else {
    for (int64_t s = plan.S - 1; s >= 0; --s) {
      const int64_t k0 = born_seg_start(plan, nts - 1, s), k1 = born_seg_start(plan, nts - 1, s + 1);
      const int64_t kend = std::min(k1 + 1, nts - 2);  // last recomputed coarse step
      base = k0;
      if (s < plan.S - 1) {  // the last segment's samples are in the window already
        // restore the background state and the partial V[k0]
        const T* c = ck + s * 21 * np;
        for (int f9 = 0; f9 < 9; ++f9) {
          HIPCHK(hipMemcpyAsync(bg.p0[f9], c + f9 * np, np * sizeof(T), hipMemcpyDeviceToDevice, h->stream));
          HIPCHK(hipMemcpyAsync(bg.p1[f9], c + (9 + f9) * np, np * sizeof(T), hipMemcpyDeviceToDevice, h->stream));
        }
        HIPCHK(hipMemcpyAsync(V(k0), c + 18 * np, 3 * np * sizeof(T), hipMemcpyDeviceToDevice, h->stream));
        CHK(fill<T>(h, V(k0 + 1), (kend + 1 - k0) * 3 * np));
        for (int64_t k = k0; k <= kend; ++k) CHK(bg_coarse_step<T>(h, bg, sp, src_tr, k, V(k), V(k + 1)));
        steps += (kend + 1 - k0) * sub;
      }
      // V[k0 .. min(k1 + 1, nts - 1)] are final
      for (int64_t its = k1 - 1; its >= k0; --its) CHK(adjoint_step(its));
    }
  }
  launch_imaging<T, true>(h, imaging_args<T>(h, V(-1), V(0), V(1), 0, nullptr, dm, ssR));
  h->last_steps = steps;
  return ELA_OK;
}

template <typename T>
int born_run(ela_handle* h, const ela_points* sp_in, const ela_traces* st_in, const void* dmod_in,
             void* dmod_out, const ela_points* rp_in, const ela_traces* rt_in, bool adjoint, int mem) {
  if (!h->have_medium) return fail(ELA_ERR_STATE, "medium not set");
  if (h->cfg.nts < 3) return fail(ELA_ERR_ARG, "Born operators need nts >= 3");
  CHK(check_points(sp_in, "source"));
  CHK(check_points(rp_in, "receiver"));
  const int64_t nts = h->cfg.nts, np = h->stride;
  DevPoints sp, rp;
  CHK(convert_points(h, sp_in, rp_in, mem, sp, rp));
  void* ds[9];
  void* dr[9];
  CHK(stage_traces(h, sp_in, st_in, rp_in, rt_in, true, adjoint, mem, ds, dr));
  T* src_tr[9];
  T* rec_tr[9];
  for (int c = 0; c < 9; ++c) { src_tr[c] = (T*)ds[c]; rec_tr[c] = (T*)dr[c]; }

  // perturbations (forward) / incoming images (adjoint, accumulated) in the library layout
  CHK(h->dmod.alloc(8 * np * sizeof(T)));
  T* dm = (T*)h->dmod.p;
  CHK(fill<T>(h, dm, 8 * np));
  CHK(copy_in_3d(h, dm, adjoint ? dmod_out : dmod_in, 8, mem));

  const int64_t nrec_tot = total_points(rp);
  CHK(h->scratch_tm.alloc(std::max<int64_t>(1, nrec_tot * nts) * sizeof(T)));
  T* tm = (T*)h->scratch_tm.p;

  if (adjoint) {
    int64_t off = 0;
    for (int c = 0; c < 9; ++c) {
      const int64_t n = rp.n[kCompGrid[c]];
      launch_transpose<T>(h, rec_tr[c], tm + off, n, nts, nts, nrec_tot);
      off += n;
    }
  }
  // the adjoint sums its image over every coarse sample in FP64 (FP32
  // handles: in an FP64 copy of the incoming image, narrowed at the end)
  double* dm64 = nullptr;
  if (adjoint) {
    if constexpr (sizeof(T) == 8) {
      dm64 = reinterpret_cast<double*>(dm);
    } else {
      CHK(h->dmod64.alloc(8 * np * sizeof(double)));
      dm64 = (double*)h->dmod64.p;
      CHK(fill<double>(h, dm64, 8 * np));
      launch(h, ela::k_fold<T>, dim3((unsigned)std::min<int64_t>((8 * np + 255) / 256, 65536)), dim3(256), dm, dm64, 8 * np);
    }
  }
  HIPCHK(hipEventRecord(h->ev0, h->stream));
  if (!adjoint) CHK(born_forward_run<T>(h, sp, src_tr, rp, tm, nrec_tot, dm));
  else CHK(born_adjoint_run<T>(h, sp, src_tr, rp, tm, nrec_tot, dm64));
  HIPCHK(hipEventRecord(h->ev1, h->stream));
  if (adjoint && sizeof(T) == 4)
    launch(h, ela::k_narrow<T>, dim3((unsigned)std::min<int64_t>((8 * np + 255) / 256, 65536)), dim3(256), dm64, dm, 8 * np);
  HIPCHK(hipGetLastError());

  if (!adjoint) {
    int64_t off = 0;
    for (int c = 0; c < 9; ++c) {
      const int64_t n = rp.n[kCompGrid[c]];
      launch_transpose<T>(h, tm + off, rec_tr[c], nts, n, nrec_tot, nts);
      off += n;
    }
    HIPCHK(hipGetLastError());
    if (mem == ELA_MEM_HOST) CHK(stage_out_traces(h, rp_in, dr, rt_in));
  } else {
    CHK(copy_out_3d(h, dmod_out, dm, 8, mem));
  }
  if (mem == ELA_MEM_HOST) HIPCHK(hipStreamSynchronize(h->stream));
  return ELA_OK;
}

// ---- adjoint-state gradient --------------------------------------------------
//
// Gradient of a data functional J with respect to the 8 scaled medium arrays
// for the nonlinear forward operator: the transpose of the linearised fused
// step  n = W (W o + M D c + s),
//   dJ/dM = sum_j mu_j (.) D(c_j),
// mu_j being the transposed recursion (the Born-adjoint stepper driven by
// dJ/d(data) injected at the receivers). The forward pass stores checkpoints
// of its state every K1 fine steps (bfloat16 or native precision). The
// backward pass walks the checkpoint segments from last to first: it
// recomputes each inner segment of K steps from its checkpoint while writing
// the step's derivative combinations D(c_j) to HBM (F_STORE), then runs the
// adjoint steps of the segment backwards with the imaging fused into the
// stepper (F_GRAD) and the injection (GRAD), accumulating FP32/native partial
// sums that are added into the FP64 gradient once per segment. With K1 > K
// the outer segment is first recomputed to store native inner checkpoints
// every K steps (two levels). The derivative of J with respect to the scaled
// source traces is extracted at the source points (as the nonlinear adjoint
// does) so that the caller can add the source-scaling term.

struct GradPlan {
  int64_t K = 20, K1 = 20, n1 = 0, n2 = 1;
  int fmt = ELA_CK_S16;
  size_t ck1_bytes = 0, ck2_bytes = 0;  // one checkpoint
  int dfmt = ELA_CK_NATIVE;             // storage of the derivative combinations (ELA_CK_*)
  size_t drv_bytes = 0;                 // the derivative combinations of one step
};

// Storage of the derivative combinations that the recomputation writes and
// the transposed steps read (ela_set_gradient_derivatives); the int16 format
// is FP32-only, so an FP64 handle always stores them exactly.
int drv_format(const ela_handle* h) { return h->esize == 4 ? h->drv_fmt : ELA_CK_NATIVE; }

size_t drv_step_bytes(const ela_handle* h, int fmt) {
  if (fmt == ELA_CK_S16) return (size_t)9 * (h->stride * 2 + h->stride / ela::DQ_BLOCK * 4);
  return (size_t)9 * h->stride * h->esize;
}

// Derivative combinations of one step: native arrays [9][stride], or int16
// [9][stride] followed by the float scales [9][stride / DQ_BLOCK].
struct DrvSlot {
  char* p = nullptr;
  int fmt = ELA_CK_NATIVE;
};

template <typename T>
void set_drv(const ela_handle* h, ela::StepArgs<T>& a, DrvSlot d, bool store) {
  const int64_t np = h->stride;
  for (int c = 0; c < 9; ++c) {
    if (d.fmt == ELA_CK_S16) {
      a.dq16[c] = reinterpret_cast<int16_t*>(d.p) + c * np;
      a.dqs[c] = reinterpret_cast<float*>(d.p + (size_t)9 * np * 2) + c * (np / ela::DQ_BLOCK);
    } else if (store) {
      a.dst[c] = reinterpret_cast<T*>(d.p) + c * np;
    } else {
      a.drv[c] = reinterpret_cast<const T*>(d.p) + c * np;
    }
  }
}

// Cells of the checkpoint packing: the grid interior (with stacked windows
// this includes the windows' FAT planes and the columns beyond a narrower
// window, which hold zeros).
int64_t interior_count(const ela_handle* h) {
  return (h->cfg.nz - 2 * ela::FAT) * (h->cfg.nx - 2 * ela::FAT) * (h->cfg.ny - 2 * ela::FAT);
}

// Bytes of one checkpoint (18 interior arrays) in a storage format.
size_t ck_bytes(int fmt, int64_t nI, int esize) {
  if (fmt == ELA_CK_BF16) return (size_t)18 * nI * 2;
  if (fmt == ELA_CK_S16) return (size_t)18 * (nI * 2 + (nI + ela::CK_BLOCK - 1) / ela::CK_BLOCK * 4);
  return (size_t)18 * nI * esize;
}

int grad_plan(ela_handle* h, GradPlan& p) {
  const int64_t N = (h->cfg.nts - 1) * h->cfg.sub;
  const int64_t nI = interior_count(h);
  p.K = std::max<int64_t>(1, std::min<int64_t>(h->grad_interval, N));
  p.fmt = h->grad_fmt;
  p.dfmt = drv_format(h);
  p.drv_bytes = drv_step_bytes(h, p.dfmt);
  p.ck1_bytes = ck_bytes(p.fmt, nI, h->esize);
  p.ck2_bytes = (size_t)18 * nI * h->esize;
  const double arr = double(h->stride) * h->esize;
  const double fixed = double(p.drv_bytes) * p.K + arr * (8 + 36) + double(h->stride) * 8.0 * 8;  // drv, g32, wave2, g64
  auto mem_for = [&](int64_t K1) {
    const int64_t n1 = (N + K1 - 1) / K1;
    return fixed + double(n1) * p.ck1_bytes + (K1 > p.K ? double(K1 / p.K) * p.ck2_bytes : 0.0);
  };
  if (h->grad_outer > 0) {
    if (h->grad_outer % p.K) return fail(ELA_ERR_ARG, "outer checkpoint interval must be a multiple of the inner one");
    p.K1 = h->grad_outer;
  } else {
    double budget;
    if (h->grad_budget > 0) {
      budget = double(h->grad_budget);
    } else {
      size_t fr = 0, tot = 0;
      HIPCHK(hipMemGetInfo(&fr, &tot));
      budget = double(fr) + double(h->ck1.bytes + h->ck2.bytes + h->drv.bytes + h->g32.bytes + h->g64.bytes +
                                   h->wave2.bytes + h->coarse.bytes + h->ckpt.bytes + h->ss.bytes) -
               double(2ull << 30);
    }
    p.K1 = 0;
    if (mem_for(p.K) <= budget) {
      p.K1 = p.K;
    } else {
      double best = 0;
      for (int64_t m = 2; m * p.K <= N + p.K; ++m) {
        const double need = mem_for(m * p.K);
        if (need <= budget && (p.K1 == 0 || need < best)) { p.K1 = m * p.K; best = need; }
      }
    }
    if (p.K1 == 0)
      return fail(ELA_ERR_MEM, "gradient: %.1f GB of device memory cannot hold the checkpoints", budget / 1e9);
  }
  p.n1 = (N + p.K1 - 1) / p.K1;
  p.n2 = p.K1 / p.K;
  return ELA_OK;
}

// Multiplier and shifts of ela::fdiv for divisor d >= 1 (d = 0, an empty
// interior, gives a divisor nothing is divided by).
ela::FastDiv fast_div(uint32_t d) {
  if (d == 0) return ela::FastDiv{0, 0, 0};
  int l = 0;
  while (l < 32 && ((uint64_t)1 << l) < d) ++l;
  ela::FastDiv f;
  f.m = (uint32_t)((((uint64_t)1 << 32) * (((uint64_t)1 << l) - d)) / d + 1);
  f.s1 = l < 1 ? l : 1;
  f.s2 = l > 1 ? l - 1 : 0;
  return f;
}

template <typename T>
ela::CkArgs<T> ck_args(ela_handle* h, const Wave<T>& w) {
  ela::CkArgs<T> a;
  for (int k = 0; k < 9; ++k) { a.f[k] = w.p0[k]; a.f[9 + k] = w.p1[k]; }
  a.nzI = (int)(h->cfg.nz - 2 * ela::FAT);
  a.nxI = (int)(h->cfg.nx - 2 * ela::FAT);
  a.nyI = (int)(h->cfg.ny - 2 * ela::FAT);
  a.nzi = (int)h->nzi;
  a.zoff = (int)h->zoff;
  a.ys = h->cfg.nx * h->nzi;
  a.nI = interior_count(h);
  a.dz = fast_div((uint32_t)a.nzI);
  a.dx = fast_div((uint32_t)a.nxI);
  return a;
}

dim3 ck_grid(int64_t nI) { return dim3((unsigned)std::min<int64_t>((nI + 255) / 256, 4096), 18); }

// The vectorised int16 checkpoint kernels: FP32, interior z extent a multiple
// of 8, checkpoint 16-byte aligned (the int16 arrays of all 18 fields then are).
bool ck_vec8(const ela_handle* h, const void* ck) {
  return h->esize == 4 && (h->cfg.nz - 2 * ela::FAT) % 8 == 0 && reinterpret_cast<uintptr_t>(ck) % 16 == 0;
}

dim3 ck_grid8(int64_t nI) { return dim3((unsigned)std::min<int64_t>((nI / 8 + 255) / 256, 4096), 18); }

template <typename T>
int save_checkpoint(ela_handle* h, const Wave<T>& w, void* out, int fmt, float* scale) {
  ela::CkArgs<T> a = ck_args<T>(h, w);
  const dim3 grid = ck_grid(a.nI);
  if (fmt == ELA_CK_S16) {
    if constexpr (sizeof(T) == 4) {
      if (ck_vec8(h, out)) {
        launch(h, ela::k_ck_pack_s16_v8, ck_grid8(a.nI), dim3(256), a, static_cast<int16_t*>(out));
        return ELA_OK;
      }
    }
    launch(h, ela::k_ck_pack_s16<T>, grid, dim3(256), a, static_cast<int16_t*>(out));
    return ELA_OK;
  }
  uint32_t* amax = reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(h->ckscale.p) + h->ckscale.bytes - 18 * 4);
  if (fmt == ELA_CK_BF16) {
    launch(h, ela::k_fill<uint32_t>, dim3(1), dim3(64), amax, (int64_t)18, 0u);
    launch(h, ela::k_ck_amax<T>, grid, dim3(256), a, amax);
    launch(h, ela::k_ck_pack<T, true>, grid, dim3(256), a, out, amax, scale);
  } else {
    launch(h, ela::k_ck_pack<T, false>, grid, dim3(256), a, out, amax, scale);
  }
  return ELA_OK;
}

template <typename T>
int load_checkpoint(ela_handle* h, const Wave<T>& w, const void* in, int fmt, const float* scale) {
  ela::CkArgs<T> a = ck_args<T>(h, w);
  if constexpr (sizeof(T) == 4) {
    if (fmt == ELA_CK_S16 && ck_vec8(h, in)) {
      launch(h, ela::k_ck_unpack_s16_v8, ck_grid8(a.nI), dim3(256), a, static_cast<const int16_t*>(in));
      return ELA_OK;
    }
  }
  if (fmt == ELA_CK_S16)
    launch(h, ela::k_ck_unpack_s16<T>, ck_grid(a.nI), dim3(256), a,
                       static_cast<const int16_t*>(in));
  else if (fmt == ELA_CK_BF16) launch(h, ela::k_ck_unpack<T, true>, ck_grid(a.nI), dim3(256), a, in, scale);
  else launch(h, ela::k_ck_unpack<T, false>, ck_grid(a.nI), dim3(256), a, in, scale);
  return ELA_OK;
}

// One forward fine step k (source injection, optional store of D(c_k)). With
// ds.k the step index is read from device memory (the counter plus ds.off)
// and the counter advanced by ds.advance after the step (graph replays); k is
// then unused.
template <typename T, bool STORE>
int grad_fwd_step(ela_handle* h, Wave<T>& w, const DevPoints& sp, T* const src_tr[9], int64_t k, DrvSlot dst,
                  DevStep ds = DevStep()) {
  const int64_t sub = h->cfg.sub, hh = sub + 1;
  const double* f = h->filt.data();
  const int its = (int)(k / sub), i2 = (int)(k % sub);  // forward it2 = i2 + 1
  ela::StepArgs<T> a = base_args<T>(h, w.p0, w.p1);
  if (STORE) {
    set_drv<T>(h, a, dst, true);
    constexpr int FQ = sizeof(T) == 4 ? ela::F_STORE | ela::F_Q16 : ela::F_STORE;
    if (dst.fmt == ELA_CK_S16) CHK((launch_step_m<T, ela::MODE_FWD, FQ, FQ>(h, a)));
    else CHK((launch_step_m<T, ela::MODE_FWD, ela::F_STORE, ela::F_STORE>(h, a)));
  } else {
    CHK((launch_step_m<T, ela::MODE_FWD, 0, 0>(h, a)));
  }
  ela::PointArgs<T> S = make_points<T>(h, sp, w.p0, src_tr, false, 0);
  if (ds.k) set_step(h, S, ds.k, ds.off, 0);
  launch_inject<T, false>(h, S, its, f[i2], f[hh + i2]);
  if (ds.advance) launch(h, ela::k_count, dim3(1), dim3(64), ds.k, ds.advance);
  w.swap();
  return ELA_OK;
}

// g64 += g32; g32 = 0 over the 8 medium arrays: FP32 over the FD interior four
// values per lane when the interior z extent allows it, else element-wise
// over the whole arrays.
template <typename T>
int fold_gradient(ela_handle* h, T* g32, double* g64) {
  const int64_t np = h->stride;
  if constexpr (sizeof(T) == 4) {
    if ((h->cfg.nz - 2 * ela::FAT) % 4 == 0) {
      Wave<T> w{};
      ela::CkArgs<T> a = ck_args<T>(h, w);
      for (int m = 0; m < 8; ++m) a.f[m] = g32 + m * np;
      launch(h, ela::k_fold_v4, dim3((unsigned)std::min<int64_t>((a.nI / 4 + 255) / 256, 4096), 8), dim3(256), a,
             g64, np);
      return ELA_OK;
    }
  }
  const int64_t n8 = 8 * np;
  launch(h, ela::k_fold<T>, dim3((unsigned)std::min<int64_t>((n8 + 255) / 256, 65536)), dim3(256), g32, g64, n8);
  return ELA_OK;
}

// Device fine-step counter of the gradient's graph replays.
int* step_counter(ela_handle* h) { return static_cast<int*>(h->counter.p) + 1; }

void set_counter(ela_handle* h, int* c, int64_t v) {
  launch(h, ela::k_fill<int>, dim3(1), dim3(64), c, (int64_t)1, (int)v);
}

// Steps k0 .. k1 - 1 of the forward recursion without derivative store,
// native checkpoints of the state saved into `ck` every K steps from k0 when
// ck is set. With graphs, pairs of steps are replayed (K even).
template <typename T>
int grad_fwd_range(ela_handle* h, Wave<T>& w, const DevPoints& sp, T* const src_tr[9], int64_t k0, int64_t k1,
                   char* ck, int64_t K, size_t ck_bytes) {
  const bool graph = h->use_graph && (!ck || K % 2 == 0);
  hipGraphExec_t ex = nullptr;
  int* kst = nullptr;
  if (graph && k1 - k0 >= 2) {
    CHK(ensure_chunks(h));  // no synchronous upload inside the capture
    kst = step_counter(h);
    auto pair = [&]() -> int {
      Wave<T> wl = w;
      CHK((grad_fwd_step<T, false>(h, wl, sp, src_tr, 0, DrvSlot(), DevStep{kst, 0, 0})));
      CHK((grad_fwd_step<T, false>(h, wl, sp, src_tr, 0, DrvSlot(), DevStep{kst, 1, 2})));
      return ELA_OK;
    };
    CHK(graph_for(h, h->graphs[G_FWD], pair, &ex));
    set_counter(h, kst, k0);
  }
  for (int64_t k = k0; k < k1;) {
    if (ck && (k - k0) % K == 0) CHK(save_checkpoint<T>(h, w, ck + ((k - k0) / K) * ck_bytes, ELA_CK_NATIVE, nullptr));
    if (ex && k + 2 <= k1) {
      const hipError_t e = hipGraphLaunch(ex, h->stream);
      if (e != hipSuccess) return fail(ELA_ERR_HIP, "graph launch failed: %s", hipGetErrorString(e));
      k += 2;
    } else {
      CHK((grad_fwd_step<T, false>(h, w, sp, src_tr, k, DrvSlot())));
      ++k;
    }
  }
  if (ex) HIPCHK(hipEventRecord(h->graphs[G_FWD].done, h->stream));
  return ELA_OK;
}

template <typename T>
int grad_forward(ela_handle* h, const ela_points* sp_in, const ela_traces* st_in, const ela_points* rp_in,
                 const ela_traces* rt_in, int mem) {
  if (!h->have_medium) return fail(ELA_ERR_STATE, "medium not set");
  CHK(check_points(sp_in, "source"));
  CHK(check_points(rp_in, "receiver"));
  h->grad_ready = false;
  const int64_t nts = h->cfg.nts, sub = h->cfg.sub, hh = sub + 1, N = (nts - 1) * sub;
  GradPlan plan;
  CHK(grad_plan(h, plan));
  DevPoints sp, rp;
  CHK(convert_points(h, sp_in, rp_in, mem, sp, rp));
  void* ds[9];
  void* dr[9];
  CHK(stage_traces(h, sp_in, st_in, rp_in, rt_in, true, false, mem, ds, dr));
  T* src_tr[9];
  T* rec_tr[9];
  for (int c = 0; c < 9; ++c) { src_tr[c] = (T*)ds[c]; rec_tr[c] = (T*)dr[c]; }
  const int64_t nrec_tot = total_points(rp);
  CHK(h->scratch_tm.alloc(std::max<int64_t>(1, nrec_tot * nts) * sizeof(T)));
  T* tm = (T*)h->scratch_tm.p;

  if (h->ck1.bytes != plan.n1 * plan.ck1_bytes) h->ck1.release();
  CHK(h->ck1.alloc(plan.n1 * plan.ck1_bytes));
  CHK(h->ckscale.alloc((size_t)(plan.n1 + 1) * 18 * 4));
  Wave<T> w = wave_ptrs<T>(h->wave, h->stride);
  CHK(fill<T>(h, h->wave.p, 18 * h->stride));
  CHK(fill<T>(h, tm, nrec_tot * nts));
  const double* f = h->filt.data();
  // one fine step k (host index), or the step the device counter names (graph replays)
  auto step = [&](Wave<T>& wv, int64_t k, DevStep ds) -> int {
    const int its = (int)(k / sub), i2 = (int)(k % sub);
    ela::StepArgs<T> a = base_args<T>(h, wv.p0, wv.p1);
    CHK((launch_step_m<T, ela::MODE_FWD, 0, 0>(h, a)));
    ela::PointArgs<T> S = make_points<T>(h, sp, wv.p0, src_tr, false, 0);
    ela::PointArgs<T> R = make_points<T>(h, rp, wv.p0, TmViews<T>(tm).v, true, nrec_tot);
    if (ds.k) { set_step(h, S, ds.k, ds.off, 0); set_step(h, R, ds.k, ds.off, 1); }
    launch_inject<T, false>(h, S, its, f[i2], f[hh + i2]);
    launch_record<T>(h, R, its, f[i2 + 1], f[hh + i2 + 1]);
    if (ds.advance) launch(h, ela::k_count, dim3(1), dim3(64), ds.k, ds.advance);
    wv.swap();
    return ELA_OK;
  };
  // graph replays of pairs of steps (the checkpoints fall on even steps)
  hipGraphExec_t ex = nullptr;
  if (h->use_graph && plan.K1 % 2 == 0 && N >= 2) {
    int* kst = step_counter(h);
    auto pair = [&]() -> int {
      Wave<T> wl = w;
      CHK(step(wl, 0, DevStep{kst, 0, 0}));
      CHK(step(wl, 0, DevStep{kst, 1, 2}));
      return ELA_OK;
    };
    CHK(ensure_chunks(h));
    CHK(graph_for(h, h->graphs[G_FWD_REC], pair, &ex));
    set_counter(h, kst, 0);
  }
  HIPCHK(hipEventRecord(h->ev0, h->stream));
  for (int64_t k = 0; k < N;) {
    if (k % plan.K1 == 0) {
      const int64_t s1 = k / plan.K1;
      CHK(save_checkpoint<T>(h, w, (char*)h->ck1.p + s1 * plan.ck1_bytes, plan.fmt,
                             (float*)h->ckscale.p + s1 * 18));
    }
    if (ex && k + 2 <= N) {
      const hipError_t e = hipGraphLaunch(ex, h->stream);
      if (e != hipSuccess) return fail(ELA_ERR_HIP, "graph launch failed: %s", hipGetErrorString(e));
      k += 2;
    } else {
      CHK(step(w, k, DevStep()));
      ++k;
    }
  }
  if (ex) HIPCHK(hipEventRecord(h->graphs[G_FWD_REC].done, h->stream));
  HIPCHK(hipEventRecord(h->ev1, h->stream));
  h->last_steps = N;
  int64_t off = 0;
  for (int c = 0; c < 9; ++c) {
    const int64_t n = rp.n[kCompGrid[c]];
    launch_transpose<T>(h, tm + off, rec_tr[c], nts, n, nrec_tot, nts);
    off += n;
  }
  HIPCHK(hipGetLastError());
  if (mem == ELA_MEM_HOST) {
    CHK(stage_out_traces(h, rp_in, dr, rt_in));
    HIPCHK(hipStreamSynchronize(h->stream));
  }
  h->plan_K = plan.K; h->plan_K1 = plan.K1; h->plan_n1 = plan.n1; h->plan_fmt = plan.fmt;
  h->grad_ready = true;
  return ELA_OK;
}

template <typename T>
int grad_backward(ela_handle* h, const ela_points* sp_in, const ela_traces* st_in, const ela_points* rp_in,
                  const ela_traces* at_in, double* grad8, ela_traces* gs_out, int mem) {
  if (!h->grad_ready) return fail(ELA_ERR_STATE, "no gradient forward has stored checkpoints for this medium");
  CHK(check_points(sp_in, "source"));
  CHK(check_points(rp_in, "receiver"));
  const int64_t nts = h->cfg.nts, sub = h->cfg.sub, hh = sub + 1, N = (nts - 1) * sub, np = h->stride;
  GradPlan plan;
  plan.K = h->plan_K; plan.K1 = h->plan_K1; plan.n1 = h->plan_n1; plan.n2 = plan.K1 / plan.K;
  plan.fmt = h->plan_fmt;
  const int64_t nI = interior_count(h);
  plan.ck1_bytes = ck_bytes(plan.fmt, nI, h->esize);
  plan.ck2_bytes = (size_t)18 * nI * h->esize;
  plan.dfmt = drv_format(h);  // unchanged since the forward (setting it requires a new forward)
  plan.drv_bytes = drv_step_bytes(h, plan.dfmt);

  DevPoints sp, rp;
  CHK(convert_points(h, sp_in, rp_in, mem, sp, rp));
  void* ds[9];
  void* dr[9];
  CHK(stage_traces(h, sp_in, st_in, rp_in, at_in, true, true, mem, ds, dr));
  T* src_tr[9];
  T* adj_tr[9];
  for (int c = 0; c < 9; ++c) { src_tr[c] = (T*)ds[c]; adj_tr[c] = (T*)dr[c]; }
  // source-trace gradient outputs: the caller's arrays or host-call staging
  T* gs_tr[9];
  if (mem == ELA_MEM_DEVICE) {
    for (int c = 0; c < 9; ++c) {
      if (sp_in->n[kCompGrid[c]]) CHK(check_dev_ptr(gs_out->tr[c], "a source-gradient trace array"));
      gs_tr[c] = (T*)gs_out->tr[c];
    }
  } else {
    CHK(h->stage2.alloc(trace_elems(sp_in, nts) * h->esize + 9 * 256));
    size_t off = 0;
    for (int c = 0; c < 9; ++c) {
      gs_tr[c] = (T*)((char*)h->stage2.p + off);
      off += round_up(sp_in->n[kCompGrid[c]] * nts * h->esize, 256);
    }
  }
  const int64_t nrec_tot = total_points(rp), nsrc_tot = total_points(sp);
  CHK(h->scratch_tm.alloc(std::max<int64_t>(1, (nrec_tot + nsrc_tot) * nts) * sizeof(T)));
  T* tm_adj = (T*)h->scratch_tm.p;
  T* tm_src = tm_adj + nrec_tot * nts;
  int64_t off = 0;
  for (int c = 0; c < 9; ++c) {
    const int64_t n = rp.n[kCompGrid[c]];
    launch_transpose<T>(h, adj_tr[c], tm_adj + off, n, nts, nts, nrec_tot);
    off += n;
  }
  CHK(fill<T>(h, tm_src, nsrc_tot * nts));

  if (plan.n2 > 1) {
    if (h->ck2.bytes != plan.n2 * plan.ck2_bytes) h->ck2.release();
    CHK(h->ck2.alloc(plan.n2 * plan.ck2_bytes));
  }
  CHK(h->drv.alloc((size_t)plan.K * plan.drv_bytes));
  CHK(h->g32.alloc((size_t)8 * np * sizeof(T)));
  CHK(h->g64.alloc((size_t)8 * np * sizeof(double)));
  CHK(h->wave2.alloc((size_t)36 * np * sizeof(T)));
  // derivative combinations of inner step kk of a segment
  auto drv = [&](int64_t kk) { return DrvSlot{(char*)h->drv.p + kk * plan.drv_bytes, plan.dfmt}; };
  T* g32 = (T*)h->g32.p;
  double* g64 = (double*)h->g64.p;
  CHK(fill<T>(h, g32, 8 * np));
  CHK(fill<double>(h, g64, 8 * np));
  CHK(copy_in_3d(h, g64, grad8, 8, mem, sizeof(double)));
  CHK(fill<T>(h, h->wave2.p, 36 * np));
  Wave<T> bg = wave_ptrs<T>(h->wave, np), ad = wave_ptrs<T>(h->wave2, np, true);
  const double* f = h->filt.data();
  const float* scales = (const float*)h->ckscale.p;
  int64_t steps = 0;

  // adjoint step k (host index), or the step the device counter names (graph replays)
  auto adj_step = [&](Wave<T>& av, int64_t k, DrvSlot d, DevStep ds) -> int {
    const int its = (int)(k / sub), i2 = (int)(k % sub);
    ela::StepArgs<T> a = base_args<T>(h, av.p0, av.p1);
    set_pm<T>(h, a, av);
    set_drv<T>(h, a, d, false);
    for (int m = 0; m < 8; ++m) a.grad[m] = g32 + m * np;
    constexpr int FQ = sizeof(T) == 4 ? ela::F_GRAD | ela::F_Q16 : ela::F_GRAD;
    if (d.fmt == ELA_CK_S16) CHK((launch_step_m<T, ela::MODE_BORN_ADJ_PM, FQ, FQ>(h, a)));
    else CHK((launch_step_m<T, ela::MODE_BORN_ADJ_PM, ela::F_GRAD, ela::F_GRAD>(h, a)));
    ela::PointArgs<T> R = make_points<T>(h, rp, av.p0, TmViews<T>(tm_adj).v, true, nrec_tot);
    set_pm<T>(h, R, av);
    for (int c = 0; c < 9; ++c) { R.drv[c] = a.drv[c]; R.dq16[c] = a.dq16[c]; R.dqs[c] = a.dqs[c]; }
    for (int m = 0; m < 8; ++m) R.grad[m] = a.grad[m];
    ela::PointArgs<T> S = make_points<T>(h, sp, av.p0, TmViews<T>(tm_src).v, true, nsrc_tot);
    if (ds.k) { set_step(h, R, ds.k, ds.off, 1); set_step(h, S, ds.k, ds.off, 0); }
    launch_inject<T, false, true, true>(h, R, its, f[i2 + 1], f[hh + i2 + 1]);
    launch_record<T>(h, S, its, f[i2], f[hh + i2]);
    if (ds.advance) launch(h, ela::k_count, dim3(1), dim3(64), ds.k, ds.advance);
    av.swap();
    return ELA_OK;
  };
  // graph replays of whole inner segments (K steps; K even so that the
  // wavefield pointers are back in place after a segment)
  const bool seg_graph = h->use_graph && plan.K % 2 == 0;
  int* kst = step_counter(h);
  if (seg_graph) CHK(ensure_chunks(h));

  HIPCHK(hipEventRecord(h->ev0, h->stream));
  for (int64_t s1 = plan.n1 - 1; s1 >= 0; --s1) {
    const int64_t a0 = s1 * plan.K1, a1 = std::min(a0 + plan.K1, N);
    const int64_t nseg = (a1 - a0 + plan.K - 1) / plan.K;
    const void* c1 = (const char*)h->ck1.p + s1 * plan.ck1_bytes;
    if (nseg > 1) {
      CHK(load_checkpoint<T>(h, bg, c1, plan.fmt, scales + s1 * 18));
      CHK(grad_fwd_range<T>(h, bg, sp, src_tr, a0, a1, (char*)h->ck2.p, plan.K, plan.ck2_bytes));
      steps += a1 - a0;
    }
    for (int64_t s2 = nseg - 1; s2 >= 0; --s2) {
      const int64_t j0 = a0 + s2 * plan.K, j1 = std::min(j0 + plan.K, a1);
      if (nseg > 1)
        CHK(load_checkpoint<T>(h, bg, (const char*)h->ck2.p + s2 * plan.ck2_bytes, ELA_CK_NATIVE, nullptr));
      else CHK(load_checkpoint<T>(h, bg, c1, plan.fmt, scales + s1 * 18));
      if (seg_graph && j1 - j0 == plan.K) {
        // recompute with derivative store, then the transposed steps, each one replayed graph
        // the counter holds the segment's first step j0; step j0 + kk reads it at offset kk
        auto store_seg = [&]() -> int {
          Wave<T> wl = bg;
          for (int64_t kk = 0; kk < plan.K; ++kk)
            CHK((grad_fwd_step<T, true>(h, wl, sp, src_tr, 0, drv(kk), DevStep{kst, (int)kk, 0})));
          return ELA_OK;
        };
        auto adj_seg = [&]() -> int {
          Wave<T> wl = ad;
          for (int64_t kk = plan.K - 1; kk >= 0; --kk)
            CHK(adj_step(wl, 0, drv(kk), DevStep{kst, (int)kk, 0}));
          return ELA_OK;
        };
        hipGraphExec_t ex = nullptr;
        set_counter(h, kst, j0);
        CHK(graph_for(h, h->graphs[G_STORE], store_seg, &ex));
        CHK(replay(h, h->graphs[G_STORE], ex, 1));
        CHK(graph_for(h, h->graphs[G_ADJ], adj_seg, &ex));
        CHK(replay(h, h->graphs[G_ADJ], ex, 1));
      } else {
        for (int64_t k = j0; k < j1; ++k)
          CHK((grad_fwd_step<T, true>(h, bg, sp, src_tr, k, drv(k - j0))));
        for (int64_t k = j1 - 1; k >= j0; --k) CHK(adj_step(ad, k, drv(k - j0), DevStep()));
      }
      steps += 2 * (j1 - j0);
      CHK(fold_gradient<T>(h, g32, g64));
    }
  }
  HIPCHK(hipEventRecord(h->ev1, h->stream));
  h->last_steps = steps;
  CHK(copy_out_3d(h, grad8, g64, 8, mem, sizeof(double)));
  off = 0;
  for (int c = 0; c < 9; ++c) {
    const int64_t n = sp.n[kCompGrid[c]];
    launch_transpose<T>(h, tm_src + off, gs_tr[c], nts, n, nsrc_tot, nts);
    off += n;
  }
  HIPCHK(hipGetLastError());
  if (mem == ELA_MEM_HOST) {
    for (int c = 0; c < 9; ++c) {
      const int64_t n = sp_in->n[kCompGrid[c]] * nts;
      if (n) HIPCHK(hipMemcpyAsync(gs_out->tr[c], gs_tr[c], n * h->esize, hipMemcpyDeviceToHost, h->stream));
    }
    HIPCHK(hipStreamSynchronize(h->stream));
  }
  return ELA_OK;
}

template <typename T>
int time_kernels(ela_handle* h, int reps, double* ms_v, double* ms_s) {
  Wave<T> w = wave_ptrs<T>(h->wave, h->stride);
  ela::StepArgs<T> a = base_args<T>(h, w.p0, w.p1);
  float tv = 0.f, ts = 0.f;
  CHK((launch_step_m<T, ela::MODE_FWD, 0, 0>(h, a, 3)));  // warm caches and code
  HIPCHK(hipEventRecord(h->ev0, h->stream));
  for (int r = 0; r < reps; ++r) CHK((launch_step_m<T, ela::MODE_FWD, 0, 0>(h, a, 1)));
  HIPCHK(hipEventRecord(h->ev1, h->stream));
  HIPCHK(hipEventSynchronize(h->ev1));
  HIPCHK(hipEventElapsedTime(&tv, h->ev0, h->ev1));
  HIPCHK(hipEventRecord(h->ev0, h->stream));
  for (int r = 0; r < reps; ++r) CHK((launch_step_m<T, ela::MODE_FWD, 0, 0>(h, a, 2)));
  HIPCHK(hipEventRecord(h->ev1, h->stream));
  HIPCHK(hipEventSynchronize(h->ev1));
  HIPCHK(hipEventElapsedTime(&ts, h->ev0, h->ev1));
  HIPCHK(hipGetLastError());
  *ms_v = tv / reps;
  *ms_s = ts / reps;
  return ELA_OK;
}

// Mean durations of the three kernel pairs of a gradient step on the
// handle's grid and launch shape: the forward step, the recomputing step
// with derivative store (in the handle's derivative format) and the
// transposed step with imaging. Runs on the handle's wavefields and the
// gradient's own buffers (sized here for at least one step, zero-filled, and
// kept, as a gradient keeps them; the storage plans count them as free),
// whose contents do not matter for timing; stored checkpoints are untouched.
template <typename T>
int time_gradient_kernels(ela_handle* h, int reps, double* ms3) {
  const int64_t np = h->stride;
  const int dfmt = drv_format(h);
  CHK(h->wave2.alloc((size_t)36 * np * sizeof(T)));
  CHK(h->drv.alloc(drv_step_bytes(h, dfmt)));
  CHK(h->g32.alloc((size_t)8 * np * sizeof(T)));
  CHK(fill<T>(h, h->wave2.p, 36 * np));
  CHK(fill<T>(h, h->g32.p, 8 * np));
  CHK(fill<float>(h, h->drv.p, (int64_t)drv_step_bytes(h, dfmt) / 4));  // a multiple of 4 bytes
  Wave<T> w = wave_ptrs<T>(h->wave, np), ad = wave_ptrs<T>(h->wave2, np, true);
  const DrvSlot slot{(char*)h->drv.p, dfmt};
  ela::StepArgs<T> f = base_args<T>(h, w.p0, w.p1);
  ela::StepArgs<T> r = base_args<T>(h, w.p0, w.p1);
  set_drv<T>(h, r, slot, true);
  ela::StepArgs<T> g = base_args<T>(h, ad.p0, ad.p1);
  set_pm<T>(h, g, ad);
  set_drv<T>(h, g, slot, false);
  for (int m = 0; m < 8; ++m) g.grad[m] = (T*)h->g32.p + m * np;
  constexpr int FQS = sizeof(T) == 4 ? ela::F_STORE | ela::F_Q16 : ela::F_STORE;
  constexpr int FQG = sizeof(T) == 4 ? ela::F_GRAD | ela::F_Q16 : ela::F_GRAD;
  auto step = [&](int which) -> int {
    if (which == 0) return launch_step_m<T, ela::MODE_FWD, 0, 0>(h, f);
    if (which == 1) {
      if (dfmt == ELA_CK_S16) return launch_step_m<T, ela::MODE_FWD, FQS, FQS>(h, r);
      return launch_step_m<T, ela::MODE_FWD, ela::F_STORE, ela::F_STORE>(h, r);
    }
    if (dfmt == ELA_CK_S16) return launch_step_m<T, ela::MODE_BORN_ADJ_PM, FQG, FQG>(h, g);
    return launch_step_m<T, ela::MODE_BORN_ADJ_PM, ela::F_GRAD, ela::F_GRAD>(h, g);
  };
  for (int k = 0; k < 3; ++k) {
    CHK(step(k));  // warm caches and code
    HIPCHK(hipEventRecord(h->ev0, h->stream));
    for (int i = 0; i < reps; ++i) CHK(step(k));
    HIPCHK(hipEventRecord(h->ev1, h->stream));
    HIPCHK(hipEventSynchronize(h->ev1));
    float t = 0.f;
    HIPCHK(hipEventElapsedTime(&t, h->ev0, h->ev1));
    ms3[k] = t / reps;
  }
  HIPCHK(hipGetLastError());
  return ELA_OK;
}

struct DeviceGuard {
  int prev = -1;
  explicit DeviceGuard(int dev) {
    if (hipGetDevice(&prev) != hipSuccess) prev = -1;
    if (prev != dev) (void)hipSetDevice(dev);
  }
  ~DeviceGuard() {
    int cur = -1;
    if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
  }
};

}  // namespace

template <typename T>
int set_elastic(ela_handle* h, const void* rlm3, double dtw, int mem) {
  CHK(h->elastic.alloc((size_t)3 * h->stride * h->esize));
  CHK(fill<T>(h, h->elastic.p, 3 * h->stride));
  CHK(copy_in_3d(h, h->elastic.p, rlm3, 3, mem));
  ela::MedArgs<T> a;
  std::memset(&a, 0, sizeof(a));
  for (int k = 0; k < 3; ++k) a.ep[k] = arr<T>(h->elastic, h->stride, k);
  for (int k = 0; k < 8; ++k) a.m[k] = arr<T>(h->med, h->stride, k);
  a.nz = (int)h->cfg.nz; a.nx = (int)h->cfg.nx; a.ny = (int)h->cfg.ny;
  a.nzi = (int)h->nzi; a.zoff = (int)h->zoff;
  a.ys = h->cfg.nx * h->nzi;
  a.tdt = (T)(2.0 * dtw);
  a.tdt64 = 2.0 * dtw;
  const int64_t n = h->cfg.nz * h->cfg.nx * h->cfg.ny;
  launch(h, ela::k_scale_medium<T>, dim3((unsigned)std::min<int64_t>((n + 255) / 256, 1 << 20)), dim3(256), a);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(h->stream));
  h->tdt = 2.0 * dtw;
  return ELA_OK;
}

namespace {

// Every extern "C" entry runs inside guarded(): no C++ exception (a failed
// host allocation in a std::vector, a handle, a chunk table) crosses the C ABI.
template <typename F>
int guarded(F&& f) {
  try {
    return f();
  } catch (const std::bad_alloc&) {
    return fail(ELA_ERR_MEM, "host allocation failed");
  } catch (const std::exception& e) {
    return fail(ELA_ERR_ARG, "%s", e.what());
  } catch (...) {
    return fail(ELA_ERR_ARG, "unexpected exception");
  }
}

}  // namespace

extern "C" {

const char* ela_last_error(void) { return g_err.c_str(); }

const char* ela_version(void) { return "elastic_iso_3d gfx950 0.2"; }

int ela_create(const ela_config* cfg, ela_handle** out) {
  return guarded([&]() -> int {
    if (!out) return fail(ELA_ERR_ARG, "null output handle");
    *out = nullptr;
    CHK(check_cfg(cfg));
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev));
    if (cfg->device < 0 || cfg->device >= ndev)
      return fail(ELA_ERR_ARG, "device %d not present (%d devices)", cfg->device, ndev);
    DeviceGuard dg(cfg->device);
    std::unique_ptr<ela_handle> owner(new ela_handle());   // freed on every error return and on a throw
    ela_handle* h = owner.get();
    h->cfg = *cfg;
    h->esize = cfg->precision;
    h->n_model = cfg->nz * cfg->nx * cfg->ny;
    const int64_t line = 128 / h->esize;  // elements per 128-byte cache line
    h->zoff = line - ela::FAT;
    h->nzi = round_up(h->zoff + cfg->nz, line);
    h->n_int = h->nzi * cfg->nx * cfg->ny;
    // each array carries a guard past its last plane for the stencils' unpredicated loads
    h->stride = round_up(h->n_int + ela::GUARD_ROWS * h->nzi + ela::GUARD_EXTRA, 64);
    auto bail = [&](int rc) {   // the handle's buffers go with `owner`; its stream and events here
      if (h->ev0) (void)hipEventDestroy(h->ev0);
      if (h->ev1) (void)hipEventDestroy(h->ev1);
      if (h->own_stream) (void)hipStreamDestroy(h->own_stream);
      return rc;
    };
    if ((uint64_t)h->stride * (uint64_t)h->esize >= (1ull << 32))
      return bail(fail(ELA_ERR_ARG, "grid of %lld cells exceeds 4 GiB per array (32-bit offsets)", (long long)h->n_model));
    // a blocking stream: it orders with HIP's null stream, on which callers
    // (e.g. torch's default stream) produce inputs and consume outputs
    if (hipStreamCreateWithFlags(&h->own_stream, hipStreamDefault) != hipSuccess)
      return bail(fail(ELA_ERR_HIP, "stream create failed"));
    h->stream = h->own_stream;
    if (hipEventCreate(&h->ev0) != hipSuccess || hipEventCreate(&h->ev1) != hipSuccess)
      return bail(fail(ELA_ERR_HIP, "event create failed"));
    int rc = h->wave.alloc(18 * h->stride * h->esize);
    if (rc) return bail(rc);
    rc = h->med.alloc(8 * h->stride * h->esize);
    if (rc) return bail(rc);
    if (hipMemset(h->med.p, 0, 8 * h->stride * h->esize) != hipSuccess) return bail(fail(ELA_ERR_HIP, "medium clear"));
    // time filter (nonlinearPropElasticGpuFunctions_3D.cu:224-241) and taper (:248-255)
    const int64_t sub = cfg->sub, hh = sub + 1, ntw = (cfg->nts - 1) * sub + 1;
    h->filt.resize(2 * hh);
    const double sc = 1.0 / std::sqrt(double(ntw) / double(cfg->nts));
    for (int64_t i = 0; i < hh; ++i) {
      const double a = 1.0 - double(i) / double(sub);
      h->filt[i] = a * sc;
      h->filt[hh + i] = (1.0 - a) * sc;
    }
    const int64_t mp = cfg->min_pad;
    if (mp > 0) {
      std::vector<double> tap(mp);
      for (int64_t d = 0; d < mp; ++d)
        tap[d] = cfg->alpha_cos + (1.0 - cfg->alpha_cos) * std::cos(M_PI / double(mp) * double(mp - d));
      rc = h->taper.alloc(mp * h->esize);
      if (rc) return bail(rc);
      hipError_t e;
      if (h->esize == 4) {
        std::vector<float> tf(tap.begin(), tap.end());
        e = hipMemcpy(h->taper.p, tf.data(), mp * 4, hipMemcpyHostToDevice);
      } else {
        e = hipMemcpy(h->taper.p, tap.data(), mp * 8, hipMemcpyHostToDevice);
      }
      if (e != hipSuccess) return bail(fail(ELA_ERR_HIP, "taper copy failed"));
    }
    // the filter in the handle's precision and the step counters of the graph replays
    rc = h->dfilt.alloc(2 * hh * h->esize);
    if (rc) return bail(rc);
    {
      hipError_t e;
      if (h->esize == 4) {
        std::vector<float> ff(h->filt.begin(), h->filt.end());
        e = hipMemcpy(h->dfilt.p, ff.data(), ff.size() * 4, hipMemcpyHostToDevice);
      } else {
        e = hipMemcpy(h->dfilt.p, h->filt.data(), h->filt.size() * 8, hipMemcpyHostToDevice);
      }
      if (e != hipSuccess) return bail(fail(ELA_ERR_HIP, "filter copy failed"));
    }
    rc = h->counter.alloc(256);
    if (rc) return bail(rc);
    h->by_x = cfg->precision == ELA_FP64 ? 8 : 16;
    // the device address of the allocation-free entry point's scratch on this device
    rc = scratch_address(cfg->device, &h->scratch_dev);
    if (rc) return bail(rc);
    *out = owner.release();
    return ELA_OK;
  });
}

int ela_destroy(ela_handle* h) {
  return guarded([&]() -> int {
    if (!h) return ELA_OK;
    DeviceGuard dg(h->cfg.device);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    for (GraphCache& gc : h->graphs) {
      (void)drop_graph(gc);
      if (gc.done) (void)hipEventDestroy(gc.done);
    }
    if (h->ev0) (void)hipEventDestroy(h->ev0);
    if (h->ev1) (void)hipEventDestroy(h->ev1);
    if (h->cap_stream) (void)hipStreamDestroy(h->cap_stream);
    if (h->own_stream) (void)hipStreamDestroy(h->own_stream);
    delete h;
    return ELA_OK;
  });
}

int ela_set_stream(ela_handle* h, void* s) {
  return guarded([&]() -> int {
    if (!h) return fail(ELA_ERR_ARG, "null handle");
    h->stream = s ? (hipStream_t)s : h->own_stream;
    return ELA_OK;
  });
}

int ela_synchronize(ela_handle* h) {
  return guarded([&]() -> int {
    if (!h) return fail(ELA_ERR_ARG, "null handle");
    DeviceGuard dg(h->cfg.device);
    HIPCHK(hipStreamSynchronize(h->stream));
    return ELA_OK;
  });
}

int ela_set_medium(ela_handle* h, const void* med8, int mem) {
  return guarded([&]() -> int {
    if (!h || !med8) return fail(ELA_ERR_ARG, "null argument");
    DeviceGuard dg(h->cfg.device);
    CHK(copy_in_3d(h, h->med.p, med8, 8, mem));
    HIPCHK(hipStreamSynchronize(h->stream));
    h->have_medium = true;
    h->have_elastic = false;   // the steppers read these 8 arrays
    h->grad_ready = false;
    return ELA_OK;
  });
}

int ela_set_elastic(ela_handle* h, const void* rlm3, double dtw, int mem) {
  return guarded([&]() -> int {
    if (!h || !rlm3) return fail(ELA_ERR_ARG, "null argument");
    if (!(dtw > 0)) return fail(ELA_ERR_ARG, "dtw must be positive");
    DeviceGuard dg(h->cfg.device);
    h->have_medium = h->have_elastic = false;
    CHK(h->esize == 4 ? set_elastic<float>(h, rlm3, dtw, mem) : set_elastic<double>(h, rlm3, dtw, mem));
    h->have_medium = h->have_elastic = true;
    h->grad_ready = false;
    return ELA_OK;
  });
}

int ela_get_medium(ela_handle* h, void* med8, int mem) {
  return guarded([&]() -> int {
    if (!h || !med8) return fail(ELA_ERR_ARG, "null argument");
    if (!h->have_medium) return fail(ELA_ERR_STATE, "medium not set");
    DeviceGuard dg(h->cfg.device);
    CHK(copy_out_3d(h, med8, h->med.p, 8, mem));
    HIPCHK(hipStreamSynchronize(h->stream));
    return ELA_OK;
  });
}

int ela_nl_forward(ela_handle* h, const ela_points* sp, const ela_traces* st, const ela_points* rp, ela_traces* rt,
                   int mem) {
  return guarded([&]() -> int {
    if (!h || !st || !rt) return fail(ELA_ERR_ARG, "null argument");
    DeviceGuard dg(h->cfg.device);
    return h->esize == 4 ? nl_run<float>(h, sp, st, rp, rt, false, mem) : nl_run<double>(h, sp, st, rp, rt, false, mem);
  });
}

int ela_nl_forward_noalloc(ela_handle* h, const ela_points* sp, const ela_traces* st, const ela_points* rp,
                           ela_traces* rt) {
  return guarded([&]() -> int {
    if (!h || !st || !rt || !sp || !rp) return fail(ELA_ERR_ARG, "null argument");
    if (!h->scratch_dev) return fail(ELA_ERR_STATE, "scratch address not resolved");
    DeviceGuard dg(h->cfg.device);
    return h->esize == 4 ? nl_run<float>(h, sp, st, rp, rt, false, ELA_MEM_DEVICE, h->scratch_dev)
                         : nl_run<double>(h, sp, st, rp, rt, false, ELA_MEM_DEVICE, h->scratch_dev);
  });
}

int ela_nl_adjoint(ela_handle* h, const ela_points* sp, ela_traces* st, const ela_points* rp, const ela_traces* rt,
                   int mem) {
  return guarded([&]() -> int {
    if (!h || !st || !rt) return fail(ELA_ERR_ARG, "null argument");
    DeviceGuard dg(h->cfg.device);
    return h->esize == 4 ? nl_run<float>(h, sp, st, rp, rt, true, mem) : nl_run<double>(h, sp, st, rp, rt, true, mem);
  });
}

int ela_born_forward(ela_handle* h, const ela_points* sp, const ela_traces* st, const void* dmod8,
                     const ela_points* rp, ela_traces* rt, int mem) {
  return guarded([&]() -> int {
    if (!h || !st || !rt || !dmod8) return fail(ELA_ERR_ARG, "null argument");
    DeviceGuard dg(h->cfg.device);
    return h->esize == 4 ? born_run<float>(h, sp, st, dmod8, nullptr, rp, rt, false, mem)
                         : born_run<double>(h, sp, st, dmod8, nullptr, rp, rt, false, mem);
  });
}

int ela_born_adjoint(ela_handle* h, const ela_points* sp, const ela_traces* st, void* dmod8, const ela_points* rp,
                     const ela_traces* rt, int mem) {
  return guarded([&]() -> int {
    if (!h || !st || !rt || !dmod8) return fail(ELA_ERR_ARG, "null argument");
    DeviceGuard dg(h->cfg.device);
    return h->esize == 4 ? born_run<float>(h, sp, st, nullptr, dmod8, rp, rt, true, mem)
                         : born_run<double>(h, sp, st, nullptr, dmod8, rp, rt, true, mem);
  });
}

int ela_set_gradient_checkpoints(ela_handle* h, int64_t interval, int64_t outer_interval, int format,
                                 int64_t budget_bytes) {
  return guarded([&]() -> int {
    if (!h) return fail(ELA_ERR_ARG, "null handle");
    if (interval < 1) return fail(ELA_ERR_ARG, "checkpoint interval must be >= 1");
    if (outer_interval < 0 || (outer_interval > 0 && outer_interval % interval))
      return fail(ELA_ERR_ARG, "outer interval must be 0 (automatic) or a multiple of the interval");
    if (budget_bytes < 0) return fail(ELA_ERR_ARG, "negative memory budget");
    if (format != ELA_CK_NATIVE && format != ELA_CK_BF16 && format != ELA_CK_S16)
      return fail(ELA_ERR_ARG, "checkpoint format must be ELA_CK_NATIVE, ELA_CK_BF16 or ELA_CK_S16");
    h->grad_interval = interval;
    h->grad_outer = outer_interval;
    h->grad_fmt = format;
    h->grad_budget = budget_bytes;
    h->grad_ready = false;
    return ELA_OK;
  });
}

int ela_set_gradient_derivatives(ela_handle* h, int format) {
  return guarded([&]() -> int {
    if (!h) return fail(ELA_ERR_ARG, "null handle");
    if (format != ELA_CK_NATIVE && format != ELA_CK_S16)
      return fail(ELA_ERR_ARG, "derivative format must be ELA_CK_NATIVE or ELA_CK_S16");
    if (format == ELA_CK_S16 && h->esize != 4) return fail(ELA_ERR_ARG, "ELA_CK_S16 derivatives need an FP32 handle");
    h->drv_fmt = format;
    h->grad_ready = false;  // the plan of the next gradient forward sizes the derivative buffers
    return ELA_OK;
  });
}

int ela_gradient_derivatives_used(const ela_handle* h, int* format) {
  return guarded([&]() -> int {
    if (!h || !format) return fail(ELA_ERR_ARG, "null argument");
    *format = drv_format(h);
    return ELA_OK;
  });
}

int ela_gradient_plan_used(const ela_handle* h, int64_t* interval, int64_t* outer_interval, int* format,
                           int64_t* checkpoint_bytes) {
  return guarded([&]() -> int {
    if (!h || !interval || !outer_interval || !format || !checkpoint_bytes) return fail(ELA_ERR_ARG, "null argument");
    *interval = h->plan_K;
    *outer_interval = h->plan_K1;
    *format = h->plan_fmt;
    *checkpoint_bytes = (int64_t)h->ck1.bytes;
    return ELA_OK;
  });
}

int ela_gradient_forward(ela_handle* h, const ela_points* sp, const ela_traces* st, const ela_points* rp,
                         ela_traces* rt, int mem) {
  return guarded([&]() -> int {
    if (!h || !st || !rt) return fail(ELA_ERR_ARG, "null argument");
    DeviceGuard dg(h->cfg.device);
    return h->esize == 4 ? grad_forward<float>(h, sp, st, rp, rt, mem) : grad_forward<double>(h, sp, st, rp, rt, mem);
  });
}

int ela_gradient_backward(ela_handle* h, const ela_points* sp, const ela_traces* st, const ela_points* rp,
                          const ela_traces* adj, double* grad8, ela_traces* src_grad, int mem) {
  return guarded([&]() -> int {
    if (!h || !st || !adj || !grad8 || !src_grad) return fail(ELA_ERR_ARG, "null argument");
    DeviceGuard dg(h->cfg.device);
    return h->esize == 4 ? grad_backward<float>(h, sp, st, rp, adj, grad8, src_grad, mem)
                         : grad_backward<double>(h, sp, st, rp, adj, grad8, src_grad, mem);
  });
}

int ela_set_windows(ela_handle* h, int n, const int64_t* y0, const int64_t* ny, const int64_t* nx) {
  return guarded([&]() -> int {
    if (!h) return fail(ELA_ERR_ARG, "null handle");
    if (n < 0 || (n > 0 && (!y0 || !ny || !nx))) return fail(ELA_ERR_ARG, "bad window list");
    DeviceGuard dg(h->cfg.device);
    h->win_y0.clear(); h->win_ny.clear(); h->win_nx.clear();
    h->chunks_yc = 0;
    h->grad_ready = false;
    if (n == 0) return ELA_OK;
    const int64_t NY = h->cfg.ny;
    std::vector<int> tab(4 * NY, 0);
    int64_t prev_end = 0;
    for (int w = 0; w < n; ++w) {
      if (y0[w] < prev_end || ny[w] < 2 * ela::FAT + 1 || y0[w] + ny[w] > NY || nx[w] < 2 * ela::FAT + 1 ||
          nx[w] > h->cfg.nx)
        return fail(ELA_ERR_ARG, "window %d (y0 %lld, ny %lld, nx %lld) overlaps, is too small or leaves the grid", w,
                    (long long)y0[w], (long long)ny[w], (long long)nx[w]);
      if (std::min(nx[w], h->cfg.nz) / 2 < h->cfg.min_pad || ny[w] / 2 < h->cfg.min_pad)
        return fail(ELA_ERR_ARG, "window %d is narrower than twice the sponge", w);
      for (int64_t y = y0[w]; y < y0[w] + ny[w]; ++y) {
        tab[4 * y] = (int)y0[w];
        tab[4 * y + 1] = (int)ny[w];
        tab[4 * y + 2] = (int)nx[w];
      }
      prev_end = y0[w] + ny[w];
      h->win_y0.push_back(y0[w]); h->win_ny.push_back(ny[w]); h->win_nx.push_back(nx[w]);
    }
    CHK(h->ytab.alloc(tab.size() * sizeof(int)));
    HIPCHK(hipMemcpy(h->ytab.p, tab.data(), tab.size() * sizeof(int), hipMemcpyHostToDevice));
    CHK(ensure_chunks(h));
    return ELA_OK;
  });
}

int ela_set_graph(ela_handle* h, int on) {
  return guarded([&]() -> int {
    if (!h) return fail(ELA_ERR_ARG, "null handle");
    h->use_graph = on != 0;
    return ELA_OK;
  });
}

int ela_set_tiling(ela_handle* h, int tile_x, int y_chunk) {
  return guarded([&]() -> int {
    if (!h) return fail(ELA_ERR_ARG, "null handle");
    if (tile_x != 4 && tile_x != 8 && tile_x != 16) return fail(ELA_ERR_ARG, "tile_x must be 4, 8 or 16");
    if (y_chunk < 1) return fail(ELA_ERR_ARG, "y_chunk must be >= 1");
    h->by_x = tile_x;
    h->ychunk = y_chunk;
    return ELA_OK;
  });
}

int ela_set_tiling_transposed(ela_handle* h, int tile_x) {
  return guarded([&]() -> int {
    if (!h) return fail(ELA_ERR_ARG, "null handle");
    if (tile_x != 0 && tile_x != 4 && tile_x != 8 && tile_x != 16)
      return fail(ELA_ERR_ARG, "tile_x must be 0 (the forward tile), 4, 8 or 16");
    h->by_x_t = tile_x;
    return ELA_OK;
  });
}

int ela_get_tiling_transposed(const ela_handle* h, int* tile_x) {
  return guarded([&]() -> int {
    if (!h || !tile_x) return fail(ELA_ERR_ARG, "null argument");
    *tile_x = h->by_x_t ? h->by_x_t : h->by_x;
    return ELA_OK;
  });
}

int ela_set_tile_order(ela_handle* h, int order, int x_group) {
  return guarded([&]() -> int {
    if (!h) return fail(ELA_ERR_ARG, "null handle");
    if (order < 0 || order > 3) return fail(ELA_ERR_ARG, "tile order must be 0, 1, 2 or 3");
    if (x_group < 1) return fail(ELA_ERR_ARG, "x_group must be >= 1");
    h->remap = order;
    h->xgroup = x_group;
    return ELA_OK;
  });
}

int ela_set_born_storage(ela_handle* h, int64_t segment, int64_t budget_bytes) {
  return guarded([&]() -> int {
    if (!h) return fail(ELA_ERR_ARG, "null handle");
    if (segment < -1) return fail(ELA_ERR_ARG, "segment must be -1, 0 or a positive number of coarse samples");
    if (budget_bytes < 0) return fail(ELA_ERR_ARG, "negative memory budget");
    h->born_segment = segment;
    h->born_budget = budget_bytes;
    return ELA_OK;
  });
}

int ela_set_born_mode(ela_handle* h, int mode) {
  return guarded([&]() -> int {
    if (!h) return fail(ELA_ERR_ARG, "null handle");
    if (mode != ELA_BORN_REFERENCE && mode != ELA_BORN_PHYSICS)
      return fail(ELA_ERR_ARG, "Born mode must be ELA_BORN_REFERENCE or ELA_BORN_PHYSICS");
    h->born_physics = mode == ELA_BORN_PHYSICS;
    return ELA_OK;
  });
}

int ela_born_storage_used(const ela_handle* h, int64_t* segment) {
  return guarded([&]() -> int {
    if (!h || !segment) return fail(ELA_ERR_ARG, "null argument");
    *segment = h->born_last_segment;
    return ELA_OK;
  });
}

int ela_get_tiling(const ela_handle* h, int* tile_x, int* y_chunk) {
  return guarded([&]() -> int {
    if (!h || !tile_x || !y_chunk) return fail(ELA_ERR_ARG, "null argument");
    *tile_x = h->by_x;
    *y_chunk = h->ychunk;
    return ELA_OK;
  });
}

int ela_stream_probe(ela_handle* h, int reads, int writes, int64_t n, int reps, double* gbs) {
  return guarded([&]() -> int {
    if (!h || !gbs || reps < 1 || n < 4 || n % 4) return fail(ELA_ERR_ARG, "bad argument");
    void (*kern)(float4* const*, int64_t) = nullptr;
    if (reads == 1 && writes == 1) kern = ela::k_stream<1, 1>;
    else if (reads == 12 && writes == 3) kern = ela::k_stream<12, 3>;
    else if (reads == 14 && writes == 6) kern = ela::k_stream<14, 6>;
    else if (reads == 10 && writes == 3) kern = ela::k_stream<10, 3>;
    else if (reads == 11 && writes == 6) kern = ela::k_stream<11, 6>;
    else if (reads == 12 && writes == 9) kern = ela::k_stream<12, 9>;
    else if (reads == 12 && writes == 0) kern = ela::k_stream<12, 0>;
    else if (reads == 2 && writes == 0) kern = ela::k_stream<2, 0>;
    if (!kern)
      return fail(ELA_ERR_ARG, "stream probe mixes: (1, 1), (12, 3), (14, 6), (10, 3), (11, 6), (12, 9), (12, 0), (2, 0)");
    DeviceGuard dg(h->cfg.device);
    DevBuf buf, tab;
    CHK(buf.alloc((size_t)reads * n * 4));
    CHK(tab.alloc(sizeof(float4*) * 16));
    float4* ptrs[16];
    for (int k = 0; k < reads; ++k) ptrs[k] = reinterpret_cast<float4*>(static_cast<char*>(buf.p) + (size_t)k * n * 4);
    HIPCHK(hipMemcpyAsync(tab.p, ptrs, sizeof(float4*) * reads, hipMemcpyHostToDevice, h->stream));
    CHK(fill<float>(h, static_cast<float*>(buf.p), (int64_t)reads * n));
    const int64_t n4 = n / 4;
    // enough 256-thread blocks for about 4 vectors per thread
    const unsigned blocks = (unsigned)std::max<int64_t>(1, std::min<int64_t>((n4 + 1023) / 1024, 1 << 20));
    float4* const* t = static_cast<float4* const*>(tab.p);
    launch(h, kern, dim3(blocks), dim3(256), t, n4);
    HIPCHK(hipEventRecord(h->ev0, h->stream));
    for (int r = 0; r < reps; ++r) launch(h, kern, dim3(blocks), dim3(256), t, n4);
    HIPCHK(hipEventRecord(h->ev1, h->stream));
    HIPCHK(hipEventSynchronize(h->ev1));
    HIPCHK(hipGetLastError());
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, h->ev0, h->ev1));
    *gbs = double(reads + writes) * 4.0 * double(n) * reps / (double(ms) * 1e-3) / 1e9;
    HIPCHK(hipStreamSynchronize(h->stream));
    return ELA_OK;
  });
}

int ela_time_step_kernels(ela_handle* h, int reps, double* ms_velocity, double* ms_stress) {
  return guarded([&]() -> int {
    if (!h || reps < 1 || !ms_velocity || !ms_stress) return fail(ELA_ERR_ARG, "bad argument");
    if (!h->have_medium) return fail(ELA_ERR_STATE, "medium not set");
    DeviceGuard dg(h->cfg.device);
    return h->esize == 4 ? time_kernels<float>(h, reps, ms_velocity, ms_stress)
                         : time_kernels<double>(h, reps, ms_velocity, ms_stress);
  });
}

int ela_time_gradient_kernels(ela_handle* h, int reps, double* ms3) {
  return guarded([&]() -> int {
    if (!h || reps < 1 || !ms3) return fail(ELA_ERR_ARG, "bad argument");
    if (!h->have_medium) return fail(ELA_ERR_STATE, "medium not set");
    DeviceGuard dg(h->cfg.device);
    return h->esize == 4 ? time_gradient_kernels<float>(h, reps, ms3) : time_gradient_kernels<double>(h, reps, ms3);
  });
}

int ela_last_timing(const ela_handle* h, double* ms, int64_t* fine_steps, int64_t* cells) {
  return guarded([&]() -> int {
    if (!h) return fail(ELA_ERR_ARG, "null handle");
    float t = 0.f;
    DeviceGuard dg(h->cfg.device);
    HIPCHK(hipEventSynchronize(h->ev1));
    HIPCHK(hipEventElapsedTime(&t, h->ev0, h->ev1));
    if (ms) *ms = t;
    if (fine_steps) *fine_steps = h->last_steps;
    if (cells) *cells = interior_cells(h);
    return ELA_OK;
  });
}

}  // extern "C"
