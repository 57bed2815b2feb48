// Host side of libkmc_dn_amd.so: the C-ABI declared in include/kmc_dn_amd.h.
//
// Holds the device-resident registry (layouts, energy sets, batches) of one
// engine (one device, one HIP stream), builds the per-layout tables the
// kernels read, groups trajectories by layout for the batched kernel and
// launches the kernels (optionally through a captured hipGraph).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <chrono>
#include <mutex>
#include <thread>
#include <numeric>
#include <string>
#include <vector>

#include "kmc_types.h"

namespace kmc {

template <bool REC, bool STAGE, bool PHASE>
__global__ void kmc_plain_fp64_kernel(PlainParams);
template <int C, int W, bool REC, bool BT, bool PRIO, bool PH>
__global__ void kmc_batch_fp32_kernel(BatchParams);
template <int C, int W, bool REC>
__global__ void kmc_batch_pos_kernel(BatchParams);
__global__ void kmc_queue_reset_kernel(int32_t*);
__global__ void kmc_fill_zero_kernel(double*, int64_t);
__global__ void kmc_eset_scatter_kernel(const EsetDev*, const int32_t*, const int32_t*, const double*,
                                        const double*, int64_t);
size_t plain_lds_bytes(int S, int N, bool stage);
size_t batch_lds_bytes(bool bt, int C, int W, int S, int N, int P);
size_t pos_lds_bytes(int C, int W);
size_t store_lds_bytes(int W, int slots);
template <int C, int T, bool REC, int FORM, bool PH>
__global__ void kmc_team_kernel(BatchParams);
size_t team_lds_bytes(int form, int C, int T, int S, int N, int P);
template <int C>
__global__ void kmc_meanfield_kernel(MeanfieldParams);
template <int C>
__global__ void kmc_avg_current_kernel(MeanfieldParams);

// Device memory of kmc_simulate_discrete_noalloc, the entry point for callers
// that cannot allocate: every buffer its kernels use (descriptors, the layout's
// FP64 tables, record buffers, replayed uniforms and event logs) is carved
// from this one file-scope array, so that entry point calls no hipMalloc.  Each
// device holds its own copy for the life of the process; host code uses the
// address hipGetSymbolAddress gives for it (noalloc_scratch, looked up once per
// device when the engine is initialised), never the array's name.  Every call
// on a device shares that device's copy and leaves its contents behind; calls
// of one process are serialised by the engine lock, so two calls never overlap
// on it (callers that bypassed the lock would race).
constexpr size_t NOALLOC_TABLE = (size_t)KMC_NOALLOC_MAX_SITES * KMC_NOALLOC_MAX_SITES * sizeof(double);
constexpr size_t NOALLOC_ALIGN = 256;
constexpr size_t noalloc_up(size_t x) { return (x + NOALLOC_ALIGN - 1) & ~(NOALLOC_ALIGN - 1); }
// carve-up (byte offsets): descriptors, E, V, d, tc, ird, invd, traffic,
// se_out, occ_time, uniforms, events, totals
constexpr size_t NA_DESC = 0;
constexpr size_t NA_E = noalloc_up(NA_DESC + 1024);
constexpr size_t NA_V = noalloc_up(NA_E + sizeof(double) * KMC_NOALLOC_MAX_SITES);
constexpr size_t NA_D = noalloc_up(NA_V + sizeof(double) * KMC_PMAX);
constexpr size_t NA_TC = NA_D + NOALLOC_TABLE;
constexpr size_t NA_IRD = NA_TC + NOALLOC_TABLE;
constexpr size_t NA_INVD = NA_IRD + NOALLOC_TABLE;
constexpr size_t NA_TRAFFIC = NA_INVD + NOALLOC_TABLE;
constexpr size_t NA_SE = NA_TRAFFIC + NOALLOC_TABLE;
constexpr size_t NA_OT = noalloc_up(NA_SE + sizeof(double) * KMC_NMAX);
constexpr size_t NA_U = noalloc_up(NA_OT + sizeof(double) * KMC_NMAX);
constexpr size_t NA_EV = NA_U + sizeof(double) * 2 * KMC_NOALLOC_MAX_LOGGED_HOPS;
constexpr size_t NA_TOT = noalloc_up(NA_EV + sizeof(int32_t) * KMC_NOALLOC_MAX_LOGGED_HOPS);
constexpr size_t NOALLOC_BYTES = noalloc_up(NA_TOT + sizeof(double) * KMC_NOALLOC_MAX_LOGGED_HOPS);
static_assert(sizeof(LayoutDev) + sizeof(EsetDev) + sizeof(TrajDev) + sizeof(TrajOut) + 64 <= 1024,
              "descriptor block");
static_assert(NOALLOC_TABLE % NOALLOC_ALIGN == 0, "table alignment");

static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

// The layout tables kmc_layout_create builds on the host, built in scratch:
// ird[i*S+j] = (I_0*R)/d[i][j] and invd[j*N+i] = 1/d[i][j] for acceptor pairs
// i != j (0 elsewhere; the same IEEE quotients as the host's), and the record
// buffers the kernel accumulates into set to zero.
__global__ void kmc_noalloc_prep_kernel(const double* d, double* ird, double* invd, double* traffic,
                                        double* occ_time, int N, int S, double I0R, int clear) {
  const int SS = S * S;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < SS; k += gridDim.x * blockDim.x) {
    const int i = k / S, j = k - (k / S) * S;
    const bool acc = i < N && j < N && i != j;
    ird[k] = acc ? I0R / d[k] : 0.0;
    if (i < N && j < N) invd[j * N + i] = acc ? 1.0 / d[k] : 0.0;
    if (clear) traffic[k] = 0.0;
  }
  if (clear)
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < NMAX; k += gridDim.x * blockDim.x)
      occ_time[k] = 0.0;
}

namespace {

thread_local std::string g_err;

// Device address of `scratch` on each device (set when the engine initialises).
void* noalloc_scratch[64] = {};

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIPOK(expr)                                                                     \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess)                                                               \
      return fail(KMC_ERR_DEVICE, "%s failed: %s", #expr, hipGetErrorString(e_));      \
  } while (0)

struct DevBuf {
  void* p = nullptr;
  size_t n = 0;
  size_t cap = 0;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  // Contents are not preserved; an allocation that already holds `bytes` is reused.
  int alloc(size_t bytes) {
    if (p && bytes <= cap && bytes > 0) {
      n = bytes;
      return KMC_OK;
    }
    if (p) (void)hipFree(p);
    p = nullptr;
    n = cap = bytes;
    if (bytes == 0) return KMC_OK;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
      p = nullptr;
      n = cap = 0;
      return fail(KMC_ERR_NOMEM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
    }
    return KMC_OK;
  }
  template <class T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

struct Layout {
  int N = 0, P = 0, S = 0;
  std::vector<DevBuf*> eset_bufs;   // energy-set buffers created on this layout
  std::vector<int> eset_live;       // live energy sets per buffer
  std::vector<int32_t> eset_ids;    // energy-set id -> (position here)
  std::vector<int> eset_buf;        // buffer index of each entry of eset_ids
  ~Layout() {
    for (DevBuf* b : eset_bufs) delete b;
  }
  DevBuf tab32, tabB, ccol, elw, coul64, tc64, ird64, invd64, dist64, pos4;
  LayoutDev dev{};
};

struct Batch {
  int64_t ntraj = 0;
  int64_t capacity = 0;
  int group_mode = -1;              // waves per group of the current groups (0: one per group)
  std::vector<TrajDev> host;        // in layout-grouped order
  std::vector<int64_t> perm;        // grouped position -> caller index
  int smax = 0, nmax = 0, pmax = 0;
  DevBuf trajs, out, occ_time, traffic, queue, groups;
  int32_t ngroups = 0;
  bool last_rec = false;       // the last run recorded occupation time (record bit 0)
  bool last_traffic = false;   // the last run was an FP32 run with traffic (record bit 1)
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipGraphExec_t graph = nullptr;
  // what the captured graph was built from: the run options and the registry
  // generation (the device addresses of the layout and energy-set arrays)
  kmc_run_opts graph_opts{};
  uint64_t graph_gen = 0;
  // size-bucket plan: per bucket its groups, largest S / N, and a side stream
  // with fork/join events so that the bucket kernels run side by side
  std::vector<kmc_bucket> plan;
  struct Sub {
    DevBuf groups;
    int32_t ngroups = 0;
    int smax = 0, nmax = 0, pmax = 0;
  };
  Sub subs[KMC_MAX_BUCKETS];
  bool subs_built = false;
  hipStream_t side[KMC_MAX_BUCKETS] = {};
  hipEvent_t fork = nullptr, join[KMC_MAX_BUCKETS] = {};
  ~Batch() {
    if (graph) (void)hipGraphExecDestroy(graph);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (fork) (void)hipEventDestroy(fork);
    for (int b = 0; b < KMC_MAX_BUCKETS; ++b) {
      if (join[b]) (void)hipEventDestroy(join[b]);
      if (side[b]) (void)hipStreamDestroy(side[b]);
    }
  }
};

struct Engine {
  std::mutex mu;
  bool ready = false;
  int device = -1;
  int cu_count = 0;
  int lds_bytes = 0;
  int wall_khz = 0;   // rate of the kernels' constant-rate clock (wall_clock64)
  char name[256] = {0};
  hipStream_t stream = nullptr;
  std::vector<Layout*> layouts;
  std::vector<EsetDev> esets_host;
  std::vector<int> eset_layout;
  DevBuf layouts_dev, esets_dev;
  DevBuf stage;   // staged energy sets of kmc_eset_write_many
  std::vector<Batch*> batches;
  bool layouts_dirty = true, esets_dirty = true;
  // bumped whenever layouts_dev or esets_dev moves to a new device address:
  // a captured graph holds those addresses as kernel arguments
  uint64_t reg_gen = 1;
};

Engine& eng() {
  static Engine e;
  return e;
}

int ensure_init() {
  Engine& E = eng();
  if (E.ready) return KMC_OK;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fail(KMC_ERR_DEVICE, "no HIP device: %s", hipGetErrorString(e));
  E.device = dev;
  hipDeviceProp_t prop;
  HIPOK(hipGetDeviceProperties(&prop, dev));
  E.cu_count = prop.multiProcessorCount;
  HIPOK(hipDeviceGetAttribute(&E.wall_khz, hipDeviceAttributeWallClockRate, dev));
  if (E.wall_khz <= 0) return fail(KMC_ERR_DEVICE, "no wall-clock rate");
  E.lds_bytes = (int)prop.sharedMemPerBlock;
  snprintf(E.name, sizeof E.name, "%s", prop.gcnArchName);
  HIPOK(hipStreamCreateWithFlags(&E.stream, hipStreamNonBlocking));
  // the allocation-free entry point's scratch: its address on this device
  if (dev < 0 || dev >= 64) return fail(KMC_ERR_DEVICE, "device index %d out of range", dev);
  if (!noalloc_scratch[dev]) {
    void* sp = nullptr;
    HIPOK(hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch)));
    if (!sp) return fail(KMC_ERR_DEVICE, "no device address for the scratch array");
    noalloc_scratch[dev] = sp;
  }
  E.ready = true;
  return KMC_OK;
}

int sync_registry() {
  Engine& E = eng();
  const void* lp = E.layouts_dev.p;
  const void* ep = E.esets_dev.p;
  if (E.layouts_dirty) {
    std::vector<LayoutDev> v;
    v.reserve(E.layouts.size());
    for (Layout* L : E.layouts) v.push_back(L ? L->dev : LayoutDev{});
    int st = E.layouts_dev.alloc(std::max<size_t>(1, v.size()) * sizeof(LayoutDev));
    if (E.layouts_dev.p != lp) ++E.reg_gen;
    if (st) return st;
    if (!v.empty()) HIPOK(hipMemcpy(E.layouts_dev.p, v.data(), v.size() * sizeof(LayoutDev), hipMemcpyHostToDevice));
    E.layouts_dirty = false;
  }
  if (E.esets_dirty) {
    int st = E.esets_dev.alloc(std::max<size_t>(1, E.esets_host.size()) * sizeof(EsetDev));
    if (E.esets_dev.p != ep) ++E.reg_gen;
    if (st) return st;
    if (!E.esets_host.empty())
      HIPOK(hipMemcpy(E.esets_dev.p, E.esets_host.data(), E.esets_host.size() * sizeof(EsetDev),
                      hipMemcpyHostToDevice));
    E.esets_dirty = false;
  }
  return KMC_OK;
}

template <class T>
int upload(DevBuf& b, const std::vector<T>& v) {
  int st = b.alloc(v.size() * sizeof(T));
  if (st) return st;
  if (!v.empty()) HIPOK(hipMemcpy(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  return KMC_OK;
}

// Sizes, constants and tables of one network (no allocation): S = N + P <=
// smax, kT > 0, nu > 0, a symmetric transitions_constant, distinct acceptors.
int check_network(int64_t N, int64_t P, int smax, double nu, double kT, double I_0, double R,
                  const double* d, const double* tc) {
  if (N < 0 || P < 0 || N > NMAX || P > PMAX || N + P <= 0)
    return fail(KMC_ERR_LIMIT, "network N=%lld P=%lld outside 0<=N<=%d, 0<=P<=%d",
                (long long)N, (long long)P, NMAX, PMAX);
  if (N + P > smax)
    return fail(KMC_ERR_LIMIT, "S=N+P=%lld exceeds %d sites", (long long)(N + P), smax);
  if (!d || !tc) return fail(KMC_ERR_ARGS, "null table pointer");
  if (!(kT > 0) || !(nu > 0) || !std::isfinite(I_0) || !std::isfinite(R))
    return fail(KMC_ERR_ARGS, "need kT > 0, nu > 0 and finite I_0, R");
  const int S = (int)(N + P);
  for (int i = 0; i < S; ++i)
    for (int j = 0; j < i; ++j) {
      const double a = tc[i * S + j], b = tc[j * S + i];
      if (std::fabs(a - b) > 1e-12 * std::max(std::fabs(a), std::fabs(b)))
        return fail(KMC_ERR_ARGS, "transitions_constant is not symmetric at (%d,%d)", i, j);
      if (i < N && j < N && !(d[i * S + j] > 0))
        return fail(KMC_ERR_ARGS, "acceptors %d and %d coincide (distance %g)", i, j, d[i * S + j]);
    }
  return KMC_OK;
}

int layout_create_locked(int64_t N, int64_t P, double nu, double kT, double I_0, double R,
                         const double* d, const double* tc, double prune, int32_t* out) {
  int st0 = check_network(N, P, SMAX_FP32, nu, kT, I_0, R, d, tc);
  if (st0) return st0;
  if (!out) return fail(KMC_ERR_ARGS, "null layout id pointer");
  const int S = (int)(N + P);
  const int n = (int)N;
  const double I0R = I_0 * R;
  const double scale2 = 1.4426950408889634 / kT;
  double tcmax = 0.0;
  for (int k = 0; k < S * S; ++k) tcmax = std::max(tcmax, tc[k]);
  const double cut = prune > 0 ? prune * tcmax : -1.0;
  std::vector<float2> tab((size_t)S * S + 64 * 3), tabB((size_t)S * S + 64 * 3);
  std::vector<float> ccol((size_t)n * S + 64 * 3, 0.f);
  for (auto& x : tab) x = make_float2(-INFINITY, 0.f);
  for (auto& x : tabB) x = make_float2(-INFINITY, -INFINITY);
  std::vector<double> coul((size_t)n * n, 0.0), tc64(tc, tc + (size_t)S * S),
      ird((size_t)S * S, 0.0), invd((size_t)n * n, 0.0);
  std::vector<float> elw((size_t)((P + 3) & ~3) * 192, 0.f);
  bool elw_ok = true;
  for (int e = 0; e < (int)P; ++e)
    for (int s2 = 0; s2 < n; ++s2) {
      const double t = tc[(size_t)s2 * S + n + e];
      if (t > 0 && t > cut) elw[(size_t)e * 192 + s2] = (float)(nu * t);
      elw_ok = elw_ok && std::isfinite(elw[(size_t)e * 192 + s2]);
    }
  const double lnu = std::log2(nu);
  for (int u = 0; u < S; ++u)
    for (int s = 0; s < S; ++s) {
      const bool both_el = s >= n && u >= n;
      const double t = tc[s * S + u];
      float lt = -INFINITY;
      if (s != u && !both_el && t > 0 && t > cut) lt = (float)(lnu + std::log2(t));
      float c = 0.f;
      if (s < n && u < n && s != u) c = (float)(scale2 * (I0R / d[s * S + u]));
      tab[(size_t)u * S + s] = make_float2(lt, c);
      tabB[(size_t)u * S + s] = make_float2(lt, lt + c);
      if (u < n) ccol[(size_t)u * S + s] = c;
    }
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j)
      if (i != j) {
        coul[(size_t)j * n + i] = I0R / d[i * S + j];
        ird[(size_t)i * S + j] = I_0 * R / d[i * S + j];
        invd[(size_t)j * n + i] = 1.0 / d[i * S + j];
      }
  Layout* L = new Layout();
  L->N = n;
  L->P = (int)P;
  L->S = S;
  int st;
  if ((st = upload(L->tab32, tab)) || (st = upload(L->tabB, tabB)) || (st = upload(L->ccol, ccol)) ||
      (st = upload(L->elw, elw)) ||
      (st = upload(L->coul64, coul)) || (st = upload(L->tc64, tc64)) ||
      (st = upload(L->dist64, std::vector<double>(d, d + (size_t)S * S))) ||
      (st = upload(L->ird64, ird)) || (st = upload(L->invd64, invd))) {
    delete L;
    return st;
  }
  LayoutDev& D = L->dev;
  D.N = n;
  D.P = (int)P;
  D.S = S;
  D.prune = prune > 0;
  D.nu = nu;
  D.kT = kT;
  D.I0 = I_0;
  D.R = R;
  D.I0R = I0R;
  D.scale2 = (float)scale2;
  D.elw_ok = elw_ok ? 1 : 0;
  D.tab32 = L->tab32.as<float2>();
  D.tabB = L->tabB.as<float2>();
  D.ccol = L->ccol.as<float>();
  D.elw = L->elw.as<float>();
  D.coul64 = L->coul64.as<double>();
  D.tc64 = L->tc64.as<double>();
  D.dist64 = L->dist64.as<double>();
  D.ird64 = L->ird64.as<double>();
  D.invd64 = L->invd64.as<double>();
  Engine& E = eng();
  int32_t id = -1;
  for (size_t i = 0; i < E.layouts.size(); ++i)
    if (!E.layouts[i]) {
      id = (int32_t)i;
      break;
    }
  if (id < 0) {
    id = (int32_t)E.layouts.size();
    E.layouts.push_back(nullptr);
  }
  E.layouts[id] = L;
  E.layouts_dirty = true;
  *out = id;
  return KMC_OK;
}

int eset_create_locked(int32_t layout, int64_t count, const double* Ec, const double* V,
                       int32_t* first) {
  Engine& E = eng();
  if (layout < 0 || layout >= (int)E.layouts.size() || !E.layouts[layout])
    return fail(KMC_ERR_ARGS, "bad layout id %d", layout);
  if (count <= 0 || !first) return fail(KMC_ERR_ARGS, "bad energy-set count");
  Layout* L = E.layouts[layout];
  const size_t nE = (size_t)count * L->N, nV = (size_t)count * L->P;
  if ((nE && !Ec) || (nV && !V)) return fail(KMC_ERR_ARGS, "null E_constant or V");
  for (size_t i = 0; i < nE; ++i)
    if (!std::isfinite(Ec[i])) return fail(KMC_ERR_ARGS, "E_constant[%zu] is not finite", i);
  for (size_t i = 0; i < nV; ++i)
    if (!std::isfinite(V[i])) return fail(KMC_ERR_ARGS, "V[%zu] is not finite", i);
  DevBuf* b = new DevBuf();
  int st = b->alloc((nE + nV) * sizeof(double) + 16);
  if (st) {
    delete b;
    return st;
  }
  if (nE) HIPOK(hipMemcpy(b->p, Ec, nE * sizeof(double), hipMemcpyHostToDevice));
  if (nV) HIPOK(hipMemcpy(b->as<double>() + nE, V, nV * sizeof(double), hipMemcpyHostToDevice));
  int32_t base = -1;
  if (count == 1)
    for (size_t i = 0; i < E.eset_layout.size(); ++i)
      if (E.eset_layout[i] < 0) {
        base = (int32_t)i;
        break;
      }
  if (base < 0) {
    base = (int32_t)E.esets_host.size();
    E.esets_host.resize(E.esets_host.size() + count);
    E.eset_layout.resize(E.eset_layout.size() + count, -1);
  }
  *first = base;
  for (int64_t i = 0; i < count; ++i) {
    EsetDev e;
    e.E = b->as<double>() + (size_t)i * L->N;
    e.V = b->as<double>() + nE + (size_t)i * L->P;
    L->eset_ids.push_back(base + (int32_t)i);
    L->eset_buf.push_back((int)L->eset_bufs.size());
    E.esets_host[base + i] = e;
    E.eset_layout[base + i] = layout;
  }
  L->eset_bufs.push_back(b);
  L->eset_live.push_back((int)count);
  E.esets_dirty = true;
  return KMC_OK;
}

// Positions form: verify the layout's tables are the reference closed forms of
// these coordinates, then store float4 coordinates and the closed-form constants.
int layout_set_positions_locked(Layout* L, double ab, const double* xyz) {
  const int S = L->S, n = L->N;
  if (L->dev.prune)   // the positions form evaluates every pair from the coordinates
    return fail(KMC_ERR_UNSUPPORTED, "the positions form does not apply a pruned pair set; "
                                     "use the table forms for pruned layouts");
  std::vector<double> d((size_t)S * S), tc((size_t)S * S);
  HIPOK(hipMemcpy(tc.data(), L->tc64.p, sizeof(double) * S * S, hipMemcpyDeviceToHost));
  const double nu = L->dev.nu;
  for (int i = 0; i < S; ++i)
    for (int j = 0; j < S; ++j) {
      const double dx = xyz[3 * i] - xyz[3 * j], dy = xyz[3 * i + 1] - xyz[3 * j + 1],
                   dz = xyz[3 * i + 2] - xyz[3 * j + 2];
      const double dd = std::sqrt(dx * dx + dy * dy + dz * dz);
      if (i == j) continue;
      if (i >= n && j >= n) continue;   // electrode pairs never hop
      const double want = nu * std::exp(-2 * dd / ab);
      const double got = tc[(size_t)i * S + j];
      if (std::fabs(got - want) > 1e-9 * want + 1e-300)
        return fail(KMC_ERR_ARGS, "transitions_constant[%d][%d]=%g is not nu*exp(-2d/ab)=%g for "
                    "these positions", i, j, got, want);
      if (i < n && j < n && !(dd > 0)) return fail(KMC_ERR_ARGS, "coincident acceptors %d, %d", i, j);
    }
  std::vector<float4> p4(S);
  for (int i = 0; i < S; ++i)
    p4[i] = make_float4((float)xyz[3 * i], (float)xyz[3 * i + 1], (float)xyz[3 * i + 2],
                        i < n ? 1.f : 0.f);
  int st = upload(L->pos4, p4);
  if (st) return st;
  const double scale2 = (double)L->dev.scale2;
  L->dev.pos4 = L->pos4.as<float4>();
  L->dev.pos_l0 = (float)(2.0 * std::log2(nu));
  L->dev.pos_mkl = (float)(-2.0 * 1.4426950408889634 / ab);
  L->dev.pos_kc = (float)(scale2 * L->dev.I0R);
  L->dev.has_pos = 1;
  eng().layouts_dirty = true;
  return KMC_OK;
}

int layout_destroy_locked(int32_t layout) {
  Engine& E = eng();
  if (layout < 0 || layout >= (int)E.layouts.size() || !E.layouts[layout])
    return fail(KMC_ERR_ARGS, "bad layout id %d", layout);
  (void)hipStreamSynchronize(E.stream);
  Layout* L = E.layouts[layout];
  for (int32_t id : L->eset_ids) {
    E.esets_host[id] = EsetDev{nullptr, nullptr};
    E.eset_layout[id] = -1;
  }
  delete L;
  E.layouts[layout] = nullptr;
  E.layouts_dirty = E.esets_dirty = true;
  return KMC_OK;
}

int eset_destroy_locked(int32_t id) {
  Engine& E = eng();
  if (id < 0 || id >= (int)E.eset_layout.size() || E.eset_layout[id] < 0)
    return fail(KMC_ERR_ARGS, "bad energy-set id %d", id);
  Layout* L = E.layouts[E.eset_layout[id]];
  for (size_t k = 0; k < L->eset_ids.size(); ++k) {
    if (L->eset_ids[k] != id) continue;
    const int bi = L->eset_buf[k];
    L->eset_ids.erase(L->eset_ids.begin() + k);
    L->eset_buf.erase(L->eset_buf.begin() + k);
    if (--L->eset_live[bi] == 0) {
      (void)hipStreamSynchronize(E.stream);
      delete L->eset_bufs[bi];
      L->eset_bufs[bi] = nullptr;
    }
    break;
  }
  E.esets_host[id] = EsetDev{nullptr, nullptr};
  E.eset_layout[id] = -1;
  E.esets_dirty = true;
  return KMC_OK;
}

int eset_write_locked(int32_t id, const double* Ec, const double* V) {
  Engine& E = eng();
  if (id < 0 || id >= (int)E.eset_layout.size() || E.eset_layout[id] < 0)
    return fail(KMC_ERR_ARGS, "bad energy-set id %d", id);
  const Layout* L = E.layouts[E.eset_layout[id]];
  if ((L->N && !Ec) || (L->P && !V)) return fail(KMC_ERR_ARGS, "null E_constant or V");
  for (int i = 0; i < L->N; ++i)
    if (!std::isfinite(Ec[i])) return fail(KMC_ERR_ARGS, "E_constant[%d] is not finite", i);
  for (int i = 0; i < L->P; ++i)
    if (!std::isfinite(V[i])) return fail(KMC_ERR_ARGS, "V[%d] is not finite", i);
  const EsetDev& e = E.esets_host[id];
  if (L->N) HIPOK(hipMemcpy(const_cast<double*>(e.E), Ec, sizeof(double) * L->N, hipMemcpyHostToDevice));
  if (L->P) HIPOK(hipMemcpy(const_cast<double*>(e.V), V, sizeof(double) * L->P, hipMemcpyHostToDevice));
  return KMC_OK;
}

int eset_write_many_locked(int64_t n, const int32_t* ids, const double* Ec, const double* V) {
  Engine& E = eng();
  if (n == 0) return KMC_OK;
  if (n < 0 || n > (1 << 24) || !ids || !Ec || !V) return fail(KMC_ERR_ARGS, "bad energy-set list");
  std::vector<int32_t> head((size_t)3 * n);   // ids, then (N, P) pairs
  for (int64_t i = 0; i < n; ++i) {
    const int32_t id = ids[i];
    if (id < 0 || id >= (int)E.eset_layout.size() || E.eset_layout[id] < 0)
      return fail(KMC_ERR_ARGS, "entry %lld: bad energy-set id %d", (long long)i, id);
    const Layout* L = E.layouts[E.eset_layout[id]];
    for (int k = 0; k < L->N; ++k)
      if (!std::isfinite(Ec[i * NMAX + k]))
        return fail(KMC_ERR_ARGS, "entry %lld: E_constant[%d] is not finite", (long long)i, k);
    for (int k = 0; k < L->P; ++k)
      if (!std::isfinite(V[i * PMAX + k]))
        return fail(KMC_ERR_ARGS, "entry %lld: V[%d] is not finite", (long long)i, k);
    head[i] = id;
    head[n + 2 * i] = L->N;
    head[n + 2 * i + 1] = L->P;
  }
  int st = sync_registry();
  if (st) return st;
  const size_t hb = ((sizeof(int32_t) * 3 * n + 255) / 256) * 256;
  const size_t eb = sizeof(double) * NMAX * n, vb = sizeof(double) * PMAX * n;
  HIPOK(hipStreamSynchronize(E.stream));   // the previous scatter has read the staging buffer
  if ((st = E.stage.alloc(hb + eb + vb))) return st;
  char* d = E.stage.as<char>();
  HIPOK(hipMemcpy(d, head.data(), sizeof(int32_t) * 3 * n, hipMemcpyHostToDevice));
  HIPOK(hipMemcpy(d + hb, Ec, eb, hipMemcpyHostToDevice));
  HIPOK(hipMemcpy(d + hb + eb, V, vb, hipMemcpyHostToDevice));
  const EsetDev* es = E.esets_dev.as<EsetDev>();
  const int32_t* di = reinterpret_cast<const int32_t*>(d);
  const int32_t* dn = di + n;
  const double* dE = reinterpret_cast<const double*>(d + hb);
  const double* dV = reinterpret_cast<const double*>(d + hb + eb);
  void* args[] = {&es, &di, &dn, &dE, &dV, &n};
  HIPOK(hipLaunchKernel((const void*)kmc_eset_scatter_kernel, dim3((unsigned)n), dim3(64), args, 0,
                        E.stream));
  return KMC_OK;
}

Batch* get_batch(int32_t id) {
  Engine& E = eng();
  if (id < 0 || id >= (int)E.batches.size()) return nullptr;
  return E.batches[id];
}

// Validates descriptors and fills B->host / B->perm / B->smax / B->nmax.
int batch_prepare(Batch* B, const kmc_traj* t, int64_t ntraj) {
  Engine& E = eng();
  if (!t || ntraj <= 0) return fail(KMC_ERR_ARGS, "empty batch");
  if (ntraj > (1 << 30)) return fail(KMC_ERR_LIMIT, "too many trajectories");
  for (int64_t i = 0; i < ntraj; ++i) {
    const kmc_traj& x = t[i];
    if (x.layout < 0 || x.layout >= (int)E.layouts.size() || !E.layouts[x.layout] || x.eset < 0 ||
        x.eset >= (int)E.esets_host.size() || E.eset_layout[x.eset] != x.layout)
      return fail(KMC_ERR_ARGS, "trajectory %lld: bad layout/energy-set pair (%d, %d)",
                  (long long)i, x.layout, x.eset);
    if (x.hops < 0 || x.prehops < 0 || x.hops + x.prehops > (int64_t)INT32_MAX)
      return fail(KMC_ERR_ARGS, "trajectory %lld: hops out of range", (long long)i);
    const int N = E.layouts[x.layout]->N;
    for (int w = 0; w < OCCW; ++w) {
      const int lo = w * 64;
      uint64_t valid = N <= lo ? 0 : (N - lo >= 64 ? ~0ull : ((1ull << (N - lo)) - 1));
      if (x.occ[w] & ~valid)
        return fail(KMC_ERR_ARGS, "trajectory %lld: occupation bit beyond N=%d", (long long)i, N);
    }
  }
  B->ntraj = ntraj;
  B->perm.resize(ntraj);
  std::iota(B->perm.begin(), B->perm.end(), 0);
  // group by layout, longest budgets first inside each layout
  std::stable_sort(B->perm.begin(), B->perm.end(), [&](int64_t a, int64_t b) {
    if (t[a].layout != t[b].layout) return t[a].layout < t[b].layout;
    return t[a].prehops + t[a].hops > t[b].prehops + t[b].hops;
  });
  B->host.resize(ntraj);
  B->smax = B->nmax = B->pmax = 0;
  for (int64_t i = 0; i < ntraj; ++i) {
    const kmc_traj& x = t[B->perm[i]];
    TrajDev& d = B->host[i];
    d.layout = x.layout;
    d.eset = x.eset;
    d.seed = x.seed;
    d.hop0 = x.hop0;
    d.prehops = x.prehops;
    d.hops = x.hops;
    for (int w = 0; w < OCCW; ++w) d.occ[w] = x.occ[w];
    d.h_begin = 0;
    d.h_end = x.prehops + x.hops;
    B->smax = std::max(B->smax, E.layouts[x.layout]->S);
    B->nmax = std::max(B->nmax, E.layouts[x.layout]->N);
    B->pmax = std::max(B->pmax, E.layouts[x.layout]->P);
  }
  B->group_mode = -1;
  B->subs_built = false;
  if (B->graph) {
    (void)hipGraphExecDestroy(B->graph);
    B->graph = nullptr;
  }
  return KMC_OK;
}

int batch_create_locked(const kmc_traj* t, int64_t ntraj, int32_t* out) {
  Engine& E = eng();
  if (!out) return fail(KMC_ERR_ARGS, "null batch id pointer");
  Batch* B = new Batch();
  int st = batch_prepare(B, t, ntraj);
  if (st) {
    delete B;
    return st;
  }
  B->capacity = ntraj;
  if ((st = upload(B->trajs, B->host)) || (st = B->out.alloc(sizeof(TrajOut) * ntraj)) ||
      (st = B->queue.alloc(64 * KMC_MAX_BUCKETS))) {
    delete B;
    return st;
  }
  if (hipEventCreate(&B->ev0) != hipSuccess || hipEventCreate(&B->ev1) != hipSuccess) {
    delete B;
    return fail(KMC_ERR_DEVICE, "hipEventCreate failed");
  }
  int32_t id = -1;
  for (size_t i = 0; i < E.batches.size(); ++i)
    if (!E.batches[i]) {
      id = (int32_t)i;
      break;
    }
  if (id < 0) {
    id = (int32_t)E.batches.size();
    E.batches.push_back(nullptr);
  }
  E.batches[id] = B;
  *out = id;
  return KMC_OK;
}

// The record buffers (occupation time, traffic) are sized when a run needs
// them (ensure_occ_time / ensure_traffic), against the batch's current
// capacity and largest S, so an update to more trajectories or larger
// networks never leaves them short.
int batch_update_locked(Batch* B, const kmc_traj* t, int64_t ntraj) {
  (void)hipStreamSynchronize(eng().stream);
  int st = batch_prepare(B, t, ntraj);
  if (st) return st;
  B->last_rec = B->last_traffic = false;
  if (ntraj > B->capacity) {
    if ((st = B->out.alloc(sizeof(TrajOut) * ntraj))) return st;
    B->capacity = ntraj;
    return upload(B->trajs, B->host);
  }
  HIPOK(hipMemcpy(B->trajs.p, B->host.data(), sizeof(TrajDev) * ntraj, hipMemcpyHostToDevice));
  return KMC_OK;
}

// Occupation time of every trajectory slot (NMAX doubles each), zeroed when
// (re)allocated; kernels overwrite each recorded slot at the end of a run.
int ensure_occ_time(Batch* B) {
  const size_t need = sizeof(double) * NMAX * (size_t)B->capacity;
  if (B->occ_time.p && B->occ_time.cap >= need) return KMC_OK;
  int st = B->occ_time.alloc(need);
  if (st) return st;
  HIPOK(hipMemset(B->occ_time.p, 0, need));
  return KMC_OK;
}

// Traffic matrices of every trajectory slot (smax^2 doubles each); zeroed on
// the stream by a fill kernel before each non-resume run.
int ensure_traffic(Batch* B) {
  const size_t need = sizeof(double) * (size_t)B->smax * B->smax * (size_t)B->capacity;
  if (B->traffic.p && B->traffic.cap >= need) return KMC_OK;
  return B->traffic.alloc(need);
}

int batch_set_state_locked(Batch* B, const int64_t* hb, const int64_t* he, const double* time,
                           const int64_t* eo, const uint64_t* occ) {
  Engine& E = eng();
  (void)hipStreamSynchronize(E.stream);
  if (hb || he) {
    for (int64_t i = 0; i < B->ntraj; ++i) {
      const int64_t j = B->perm[i];
      TrajDev& d = B->host[i];
      if (hb) d.h_begin = hb[j];
      if (he) d.h_end = he[j];
      if (d.h_begin < 0 || d.h_end < d.h_begin)
        return fail(KMC_ERR_ARGS, "trajectory %lld: bad hop range", (long long)j);
    }
    HIPOK(hipMemcpy(B->trajs.p, B->host.data(), sizeof(TrajDev) * B->ntraj, hipMemcpyHostToDevice));
  }
  if (time || eo || occ) {
    std::vector<TrajOut> o(B->ntraj);
    HIPOK(hipMemcpy(o.data(), B->out.p, sizeof(TrajOut) * B->ntraj, hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < B->ntraj; ++i) {
      const int64_t j = B->perm[i];
      if (time) o[i].time = time[j];
      if (eo)
        for (int e = 0; e < PMAX; ++e) o[i].eo[e] = eo[j * PMAX + e];
      if (occ)
        for (int w = 0; w < OCCW; ++w) o[i].occ[w] = occ[j * OCCW + w];
      o[i].status = KMC_OK;
    }
    HIPOK(hipMemcpy(B->out.p, o.data(), sizeof(TrajOut) * B->ntraj, hipMemcpyHostToDevice));
  }
  return KMC_OK;
}

// Groups of <= W consecutive same-layout trajectories, ordered by decreasing budget.
int build_groups(Batch* B, int W) {
  if (B->group_mode == W) return KMC_OK;
  std::vector<GroupDev> g;
  int64_t i = 0;
  while (i < B->ntraj) {
    int64_t j = i;
    while (j < B->ntraj && j - i < W && B->host[j].layout == B->host[i].layout) ++j;
    GroupDev gd;
    gd.layout = B->host[i].layout;
    gd.first = (int32_t)i;
    gd.count = (int32_t)(j - i);
    gd.pad = 0;
    g.push_back(gd);
    i = j;
  }
  std::stable_sort(g.begin(), g.end(), [&](const GroupDev& a, const GroupDev& b) {
    const TrajDev& x = B->host[a.first];
    const TrajDev& y = B->host[b.first];
    return x.prehops + x.hops > y.prehops + y.hops;
  });
  B->ngroups = (int32_t)g.size();
  B->group_mode = W;
  return upload(B->groups, g);
}

int bucket_of(const Batch* B, int S, int64_t budget) {
  for (size_t b = 0; b < B->plan.size(); ++b)
    if (S <= B->plan[b].s_max && S >= B->plan[b].s_min && budget >= B->plan[b].min_hops)
      return (int)b;
  return -1;
}

using BatchKernel = void (*)(BatchParams);

// prio: graded wave issue priority inside each hop (recording kernels always)
BatchKernel pick_batch_kernel(int C, int W, bool rec, bool bt, bool prio) {
#define KMC_PICK(CC, WW)                                                                        \
  if (C == CC && W == WW) {                                                                     \
    if (!rec && !prio)                                                                          \
      return bt ? (BatchKernel)kmc_batch_fp32_kernel<CC, WW, false, true, false, false>         \
                : (BatchKernel)kmc_batch_fp32_kernel<CC, WW, false, false, false, false>;       \
    if (bt) return rec ? (BatchKernel)kmc_batch_fp32_kernel<CC, WW, true, true, true, false>    \
                       : (BatchKernel)kmc_batch_fp32_kernel<CC, WW, false, true, true, false>;  \
    return rec ? (BatchKernel)kmc_batch_fp32_kernel<CC, WW, true, false, true, false>           \
               : (BatchKernel)kmc_batch_fp32_kernel<CC, WW, false, false, true, false>;         \
  }
  KMC_PICK(1, 1) KMC_PICK(1, 2) KMC_PICK(1, 4) KMC_PICK(1, 8)
  KMC_PICK(2, 1) KMC_PICK(2, 2) KMC_PICK(2, 4) KMC_PICK(2, 8)
  KMC_PICK(3, 1) KMC_PICK(3, 2) KMC_PICK(3, 4) KMC_PICK(3, 8)
#undef KMC_PICK
  return nullptr;
}

BatchKernel pick_team_kernel(int C, int T, bool rec, int form) {
#define KMC_PICK(CC, TT, FF)                                                         \
  if (C == CC && T == TT && form == FF)                                              \
    return rec ? (BatchKernel)kmc_team_kernel<CC, TT, true, FF, false>               \
               : (BatchKernel)kmc_team_kernel<CC, TT, false, FF, false>;
  KMC_PICK(1, 2, 1) KMC_PICK(1, 2, 2) KMC_PICK(1, 2, 3) KMC_PICK(1, 4, 1) KMC_PICK(1, 4, 2)
  KMC_PICK(1, 4, 3) KMC_PICK(2, 2, 1) KMC_PICK(2, 2, 2) KMC_PICK(2, 2, 3) KMC_PICK(2, 4, 1)
  KMC_PICK(2, 4, 2) KMC_PICK(2, 4, 3) KMC_PICK(3, 2, 1) KMC_PICK(3, 2, 2) KMC_PICK(3, 2, 3)
  KMC_PICK(3, 4, 1) KMC_PICK(3, 4, 2) KMC_PICK(3, 4, 3)
#undef KMC_PICK
  return nullptr;
}

// Phase-timing instantiations: solo table kernels of 4 waves, teams of 4 on
// two- and three-chunk networks (no records, graded issue priority).
BatchKernel pick_phase_kernel(int C, int W, int team, int form) {
  if (team == 4 && (C == 2 || C == 3) && form >= 1 && form <= 3) {
#define KMC_PICK(CC, FF) \
  if (C == CC && form == FF) return (BatchKernel)kmc_team_kernel<CC, 4, false, FF, true>;
    KMC_PICK(2, 1) KMC_PICK(2, 2) KMC_PICK(2, 3) KMC_PICK(3, 1) KMC_PICK(3, 2) KMC_PICK(3, 3)
#undef KMC_PICK
  }
  if (team <= 1 && W == 4 && (form == 1 || form == 2)) {
#define KMC_PICK(CC)                                                                  \
  if (C == CC) return form == 1 ? (BatchKernel)kmc_batch_fp32_kernel<CC, 4, false, true, true, true> \
                                : (BatchKernel)kmc_batch_fp32_kernel<CC, 4, false, false, true, true>;
    KMC_PICK(1) KMC_PICK(2) KMC_PICK(3)
#undef KMC_PICK
  }
  return nullptr;
}

BatchKernel pick_pos_kernel(int C, int W, bool rec) {
#define KMC_PICK(CC, WW)                                                             \
  if (C == CC && W == WW)                                                            \
    return rec ? (BatchKernel)kmc_batch_pos_kernel<CC, WW, true>                     \
               : (BatchKernel)kmc_batch_pos_kernel<CC, WW, false>;
  KMC_PICK(1, 1) KMC_PICK(1, 2) KMC_PICK(1, 4) KMC_PICK(1, 8)
  KMC_PICK(2, 1) KMC_PICK(2, 2) KMC_PICK(2, 4) KMC_PICK(2, 8)
  KMC_PICK(3, 1) KMC_PICK(3, 2) KMC_PICK(3, 4) KMC_PICK(3, 8)
#undef KMC_PICK
  return nullptr;
}

struct Fp32Launch {
  BatchKernel k = nullptr;
  int grid = 0, block = 0;
  size_t lds = 0;
  BatchParams p{};
};

// One work-queue kernel over `ngroups` groups.  form: 1 {lt, lt+c} + c table,
// 2 {lt, c} table, 3 positions; smax/nmax bound the S and N of its layouts.
// phases: run the phase-timing instantiation, summing into phases[FP32_PHASES].
int setup_launch(Batch* B, const kmc_run_opts& o, int form, int W, int team, int smax, int nmax,
                 int pmax, const DevBuf& groups, int32_t ngroups, int32_t* queue, int grid_cap,
                 Fp32Launch& L, uint64_t* phases) {
  Engine& E = eng();
  const int C = (smax + 63) / 64;
  // recording kernels: occupation time (bit 0) and traffic (bit 1); the others
  // carry no per-hop test for either
  const bool rec = (o.record & 3) != 0;
  const size_t lds_cap = (size_t)160 * 1024;
  BatchKernel k = nullptr;
  size_t lds = 0;
  if (team > 1) {
    k = pick_team_kernel(C, team, rec, form);
    lds = team_lds_bytes(form, C, team, smax, nmax, pmax);
    W = team;   // block of 64 * team threads, one trajectory per workgroup
  } else if (form == 3) {
    k = pick_pos_kernel(C, W, rec);
    lds = pos_lds_bytes(C, W);
  } else {
    k = pick_batch_kernel(C, W, rec, form == 1, true);
    lds = batch_lds_bytes(form == 1, C, W, smax, nmax, pmax);
  }
  if (!k) return fail(KMC_ERR_LIMIT, "no kernel for %d site chunks", C);
  if (phases) {
    if (rec) return fail(KMC_ERR_UNSUPPORTED, "phase timing runs without records");
    k = pick_phase_kernel(C, W, team, form);
    if (!k)
      return fail(KMC_ERR_UNSUPPORTED,
                  "no phase-timing kernel for %d site chunks, %d waves, team %d, table form %d "
                  "(solo: table forms with 4 waves; teams: 4 waves on 2 or 3 chunks)",
                  C, W, team, form);
  }
  if (lds > lds_cap)
    return fail(KMC_ERR_LIMIT, "LDS %zu bytes for S=%d exceeds 160 KiB (table form %d)", lds, smax, form);
  HIPOK(hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  int per_cu = 0;
  HIPOK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)k, 64 * W, lds));
  if (per_cu <= 0) return fail(KMC_ERR_LIMIT, "kernel does not fit on a CU (LDS %zu)", lds);
  // The state stores of one-chunk solo kernels: as many slots as leave the
  // resident workgroups per CU unchanged.  They change no result.  (Kernels
  // over several site chunks run their S <= 64 networks in one-chunk code
  // without a store: their LDS would take room that a plan's team
  // workgroups, placed beside them, need.)
  int slots = 0;
  if (team <= 1 && C == 1 && o.state_store >= 0) {
    const int want = o.state_store > 0 ? std::min(o.state_store, STORE_MAX_SLOTS) : STORE_MAX_SLOTS;
    for (int v = want; v > 0 && !slots; --v) {
      const size_t l2 = lds + store_lds_bytes(W, v);
      if (l2 > lds_cap) continue;
      // (refused when l2 and the kernel's static LDS exceed a workgroup's share)
      if (hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)l2) !=
          hipSuccess) {
        (void)hipGetLastError();
        continue;
      }
      int pc = 0;
      HIPOK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&pc, (const void*)k, 64 * W, l2));
      if (pc >= per_cu) {
        slots = v;
        lds = l2;
      }
    }
  }
  // The graded issue priority of the hop pays while every group is resident at
  // once (4096 trajectories: +12 %); with groups queued behind full SIMDs the
  // variant without it is 2 % faster (65,536-trajectory IV sweep).
  if (team <= 1 && form != 3 && !rec && !phases && ngroups > per_cu * E.cu_count) {
    k = pick_batch_kernel(C, W, rec, form == 1, false);
    HIPOK(hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    HIPOK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)k, 64 * W, lds));
    if (per_cu <= 0) return fail(KMC_ERR_LIMIT, "kernel does not fit on a CU (LDS %zu)", lds);
  }
  int grid = per_cu * E.cu_count;
  if (grid_cap > 0) grid = std::min(grid, grid_cap);
  const int work_units = (form == 3 && team <= 1) ? (ngroups + W - 1) / W : ngroups;
  grid = std::max(1, std::min(grid, work_units));
  BatchParams& p = L.p;
  p.trajs = B->trajs.as<TrajDev>();
  p.layouts = E.layouts_dev.as<LayoutDev>();
  p.esets = E.esets_dev.as<EsetDev>();
  p.groups = groups.as<GroupDev>();
  p.ngroups = ngroups;
  p.recompute_interval = o.recompute_interval > 0 ? o.recompute_interval : 1024;
  p.queue = queue;
  p.out = B->out.as<TrajOut>();
  p.occ_time = rec ? B->occ_time.as<double>() : nullptr;
  p.traffic = (o.record & 2) ? B->traffic.as<double>() : nullptr;
  p.traffic_stride = (int64_t)B->smax * B->smax;
  p.post_hop_record = o.post_hop_record;
  p.resume = o.resume ? 1 : 0;
  p.phases = phases;
  p.time_limit = o.time_limit_us > 0 ? (int64_t)o.time_limit_us * E.wall_khz / 1000 : 0;
  p.store_slots = slots;
  L.k = k;
  L.grid = grid;
  L.block = 64 * W;
  L.lds = lds;
  return KMC_OK;
}

// Batch-wide record buffers (occupation time, traffic).
int prepare_record(Batch* B, const kmc_run_opts& o) {
  int st;
  if ((o.record & 1) && (st = ensure_occ_time(B))) return st;
  if ((o.record & 2) && (st = ensure_traffic(B))) return st;
  return KMC_OK;
}

int check_positions(Batch* B, int bucket) {
  Engine& E = eng();
  for (const TrajDev& t : B->host)
    if ((bucket < 0 || bucket_of(B, E.layouts[t.layout]->S, t.prehops + t.hops) == bucket) &&
        !E.layouts[t.layout]->dev.has_pos)
      return fail(KMC_ERR_ARGS, "layout %d has no positions (kmc_layout_set_positions)", t.layout);
  return KMC_OK;
}

// Per-bucket groups: runs of <= W same-layout trajectories (one per group for
// the positions form), longest budgets first.
int build_bucket_groups(Batch* B) {
  if (B->subs_built) return KMC_OK;
  Engine& E = eng();
  const int nb = (int)B->plan.size();
  std::vector<std::vector<GroupDev>> g(nb);
  for (int b = 0; b < nb; ++b) B->subs[b].smax = B->subs[b].nmax = B->subs[b].pmax = 0;
  auto bucket_at = [&](int64_t k) {
    const TrajDev& t = B->host[k];
    return bucket_of(B, E.layouts[t.layout]->S, t.prehops + t.hops);
  };
  int64_t i = 0;
  while (i < B->ntraj) {
    const Layout* Ly = E.layouts[B->host[i].layout];
    const int b = bucket_at(i);
    if (b < 0)
      return fail(KMC_ERR_ARGS, "trajectory of layout %d (S=%d, budget %lld) matches no bucket",
                  B->host[i].layout, Ly->S,
                  (long long)(B->host[i].prehops + B->host[i].hops));
    const kmc_bucket& pb = B->plan[b];
    const int W = (pb.table_form == 3 || pb.team > 1) ? 1 : pb.waves_per_group;
    int64_t j = i;
    while (j < B->ntraj && j - i < W && B->host[j].layout == B->host[i].layout && bucket_at(j) == b)
      ++j;
    GroupDev gd;
    gd.layout = B->host[i].layout;
    gd.first = (int32_t)i;
    gd.count = (int32_t)(j - i);
    gd.pad = 0;
    g[b].push_back(gd);
    B->subs[b].smax = std::max(B->subs[b].smax, Ly->S);
    B->subs[b].nmax = std::max(B->subs[b].nmax, Ly->N);
    B->subs[b].pmax = std::max(B->subs[b].pmax, Ly->P);
    i = j;
  }
  for (int b = 0; b < nb; ++b) {
    std::stable_sort(g[b].begin(), g[b].end(), [&](const GroupDev& x, const GroupDev& y) {
      const TrajDev& u = B->host[x.first];
      const TrajDev& v = B->host[y.first];
      return u.prehops + u.hops > v.prehops + v.hops;
    });
    B->subs[b].ngroups = (int32_t)g[b].size();
    int st = upload(B->subs[b].groups, g[b]);
    if (st) return st;
  }
  B->subs_built = true;
  return KMC_OK;
}

// Everything that allocates or queries happens here, outside any graph capture.
int prepare_fp32(Batch* B, const kmc_run_opts& o, std::vector<Fp32Launch>& out,
                 uint64_t* phases = nullptr) {
  int st = prepare_record(B, o);
  if (st) return st;
  int32_t* q = B->queue.as<int32_t>();
  if (!B->plan.empty()) {
    if ((st = build_bucket_groups(B))) return st;
    // largest networks first: they are the slowest per hop
    for (int b = (int)B->plan.size() - 1; b >= 0; --b) {
      const Batch::Sub& sb = B->subs[b];
      if (sb.ngroups == 0) continue;
      const kmc_bucket& pb = B->plan[b];
      if (pb.table_form == 3 && (st = check_positions(B, b))) return st;
      Fp32Launch L;
      st = setup_launch(B, o, pb.table_form, pb.waves_per_group, pb.team, sb.smax, sb.nmax,
                        sb.pmax, sb.groups, sb.ngroups, q + 16 * b, pb.grid > 0 ? pb.grid : o.grid_limit, L,
                        phases);
      if (st) return st;
      out.push_back(L);
    }
    return KMC_OK;
  }
  const int C = (B->smax + 63) / 64;
  int form = o.table_form;
  int W = o.waves_per_group;
  if (W <= 0) W = form == 3 ? 4 : (C == 1 ? 4 : (C == 2 ? 2 : 1));
  if (form == 0)
    form = batch_lds_bytes(true, C, W, B->smax, B->nmax, B->pmax) <= (size_t)160 * 1024 ? 1 : 2;
  if (form < 1 || form > 3) return fail(KMC_ERR_ARGS, "table_form must be 0, 1, 2 or 3");
  if (W != 1 && W != 2 && W != 4 && W != 8) return fail(KMC_ERR_ARGS, "waves_per_group must be 1, 2, 4 or 8");
  const int team = o.team <= 1 ? 1 : o.team;
  if (team != 1 && team != 2 && team != 4) return fail(KMC_ERR_ARGS, "team must be 1, 2 or 4");
  if (form == 3 && (st = check_positions(B, -1))) return st;
  // positions and teams: one trajectory per group
  if ((st = build_groups(B, (form == 3 || team > 1) ? 1 : W))) return st;
  Fp32Launch L;
  if ((st = setup_launch(B, o, form, W, team, B->smax, B->nmax, B->pmax, B->groups, B->ngroups, q,
                         o.grid_limit, L, phases)))
    return st;
  out.push_back(L);
  return KMC_OK;
}

int launch_one(const Fp32Launch& L, hipStream_t s) {
  int32_t* q = L.p.queue;
  void* qa[] = {&q};
  HIPOK(hipLaunchKernel((const void*)kmc_queue_reset_kernel, dim3(1), dim3(1), qa, 0, s));
  BatchParams p = L.p;
  void* args[] = {&p};
  HIPOK(hipLaunchKernel((const void*)L.k, dim3(L.grid), dim3(L.block), args, L.lds, s));
  return KMC_OK;
}

// Side streams 1..n-1 (launch 0 runs on the engine stream).
int ensure_side_streams(Batch* B, int n) {
  if (!B->fork) HIPOK(hipEventCreateWithFlags(&B->fork, hipEventDisableTiming));
  for (int b = 1; b < n; ++b) {
    if (!B->side[b]) HIPOK(hipStreamCreateWithFlags(&B->side[b], hipStreamNonBlocking));
    if (!B->join[b]) HIPOK(hipEventCreateWithFlags(&B->join[b], hipEventDisableTiming));
  }
  return KMC_OK;
}

// Launches on stream s; several launches fork onto the batch's side streams
// and join back into s (parallel branches when captured into a graph).
int launch_fp32(Batch* B, const kmc_run_opts& o, const std::vector<Fp32Launch>& Ls, hipStream_t s) {
  if ((o.record & 2) && !o.resume) {
    // a kernel node, not a memset node: this may be recorded into a graph
    double* tp = B->traffic.as<double>();
    int64_t n = (int64_t)B->smax * B->smax * B->ntraj;
    void* za[] = {&tp, &n};
    const unsigned blocks = (unsigned)std::min<int64_t>(2048, (n + 255) / 256);
    HIPOK(hipLaunchKernel((const void*)kmc_fill_zero_kernel, dim3(std::max(1u, blocks)), dim3(256),
                          za, 0, s));
  }
  if (Ls.size() == 1) return launch_one(Ls[0], s);
  // launch 0 stays on s, the others fork onto side streams: n launches use n
  // streams (HIP maps streams onto a few hardware queues; launches whose
  // streams share a queue run one after the other).  They are issued in order
  // of decreasing LDS per workgroup: a workgroup that needs most of a CU's LDS
  // is placed before smaller ones take a share of every CU, which would hold it
  // back until they drain (2048-network uneven batch: 2.07 s against 3.5 s
  // with the S <= 192 team entry issued after the S <= 128 one).
  HIPOK(hipEventRecord(B->fork, s));
  std::vector<size_t> order(Ls.size());
  for (size_t b = 0; b < Ls.size(); ++b) order[b] = b;
  std::stable_sort(order.begin(), order.end(),
                   [&](size_t a, size_t b) { return Ls[a].lds > Ls[b].lds; });
  for (size_t b : order) {
    if (b == 0) {
      int st = launch_one(Ls[0], s);
      if (st) return st;
      continue;
    }
    HIPOK(hipStreamWaitEvent(B->side[b], B->fork, 0));
    int st = launch_one(Ls[b], B->side[b]);
    if (st) return st;
    HIPOK(hipEventRecord(B->join[b], B->side[b]));
  }
  for (size_t b = 1; b < Ls.size(); ++b) HIPOK(hipStreamWaitEvent(s, B->join[b], 0));
  return KMC_OK;
}

// The FP64 dense kernel over ntraj trajectories (one wave each) whose layouts
// have at most S sites and N acceptors: the layout's tables staged in LDS when
// they fit.
int launch_plain(PlainParams p, int S, int N, bool rec, int64_t ntraj, hipStream_t s) {
  const size_t cap = (size_t)160 * 1024;
  const bool stage = plain_lds_bytes(S, N, true) <= cap;
  const size_t lds = plain_lds_bytes(S, N, stage);
  if (lds > cap) return fail(KMC_ERR_LIMIT, "FP64 kernel needs %zu bytes of LDS", lds);
  using K = void (*)(PlainParams);
  K k = rec ? (stage ? (K)kmc_plain_fp64_kernel<true, true, false> : (K)kmc_plain_fp64_kernel<true, false, false>)
            : (stage ? (K)kmc_plain_fp64_kernel<false, true, false> : (K)kmc_plain_fp64_kernel<false, false, false>);
  if (p.phases) {
    if (rec) return fail(KMC_ERR_UNSUPPORTED, "phase timing runs without records");
    k = stage ? (K)kmc_plain_fp64_kernel<false, true, true> : (K)kmc_plain_fp64_kernel<false, false, true>;
  }
  HIPOK(hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  void* args[] = {&p};
  HIPOK(hipLaunchKernel((const void*)k, dim3((unsigned)ntraj), dim3(64), args, lds, s));
  return KMC_OK;
}

int enqueue_fp64(Batch* B, const kmc_run_opts& o, hipStream_t s, const double* uniforms,
                 int32_t* events, double* totals, double* traffic, double* se_out,
                 uint64_t* phases = nullptr) {
  Engine& E = eng();
  if (B->smax > SMAX_FP64)
    return fail(KMC_ERR_LIMIT, "FP64 dense kernel supports S <= %d (got %d)", SMAX_FP64, B->smax);
  const bool rec = o.record != 0 || traffic != nullptr;
  if (rec) {
    int st = ensure_occ_time(B);
    if (st) return st;
  }
  PlainParams p{};
  p.trajs = B->trajs.as<TrajDev>();
  p.layouts = E.layouts_dev.as<LayoutDev>();
  p.esets = E.esets_dev.as<EsetDev>();
  p.out = B->out.as<TrajOut>();
  p.uniforms = uniforms;
  p.events = events;
  p.totals = totals;
  p.traffic = traffic;
  p.occ_time = rec ? B->occ_time.as<double>() : nullptr;
  p.se_out = se_out;
  p.phases = phases;
  return launch_plain(p, B->smax, B->nmax, rec, B->ntraj, s);
}

int batch_run_locked(int32_t id, const kmc_run_opts* opts) {
  Engine& E = eng();
  Batch* B = get_batch(id);
  if (!B) return fail(KMC_ERR_ARGS, "bad batch id %d", id);
  kmc_run_opts o{};
  if (opts) o = *opts;
  int st = sync_registry();
  if (st) return st;
  hipStream_t s = E.stream;
  if (o.time_limit_us < 0) return fail(KMC_ERR_ARGS, "negative time limit %d us", o.time_limit_us);
  if (o.state_store < -1 || o.state_store > STORE_MAX_SLOTS)
    return fail(KMC_ERR_ARGS, "state_store %d outside [-1, %d]", o.state_store, STORE_MAX_SLOTS);
  if (o.precision == KMC_FP64_DENSE) {
    if (o.resume) return fail(KMC_ERR_UNSUPPORTED, "resume runs use the FP32 batched kernel");
    if (o.time_limit_us)
      return fail(KMC_ERR_UNSUPPORTED, "time limits are a feature of the FP32 batched kernel");
    if (o.record & 2)
      return fail(KMC_ERR_UNSUPPORTED, "batched FP64 runs record no traffic (record bit 1); "
                                       "kmc_simulate_discrete records the reference's gross traffic");
    HIPOK(hipEventRecord(B->ev0, s));
    st = enqueue_fp64(B, o, s, nullptr, nullptr, nullptr, nullptr, nullptr);
    if (st) return st;
    HIPOK(hipEventRecord(B->ev1, s));
    B->last_rec = (o.record & 1) != 0;
    B->last_traffic = false;
    return KMC_OK;
  }
  if (o.precision != KMC_FP32_BATCHED) return fail(KMC_ERR_ARGS, "unknown precision %d", o.precision);
  if (o.time_limit_us) {
    // team waves read their own clocks and could disagree on where to stop
    bool team = B->plan.empty() ? o.team > 1 : false;
    for (const kmc_bucket& pb : B->plan) team = team || pb.team > 1;
    if (team) return fail(KMC_ERR_UNSUPPORTED, "time limits apply to solo kernels only (team > 1)");
  }
  std::vector<Fp32Launch> Ls;
  if ((st = prepare_fp32(B, o, Ls))) return st;
  if (Ls.size() > 1 && (st = ensure_side_streams(B, (int)Ls.size()))) return st;
  if (o.use_graph) {
    // a graph replays its kernels' arguments as captured: re-capture when the
    // options or the registry's device addresses differ from that capture
    const bool same = B->graph && B->graph_gen == E.reg_gen &&
                      memcmp(&B->graph_opts, &o, sizeof o) == 0;
    if (!same) {
      if (B->graph) {
        (void)hipGraphExecDestroy(B->graph);
        B->graph = nullptr;
      }
      // Capture on a private stream (capture mode "relaxed" leaves other threads' work alone).
      hipStream_t cs;
      HIPOK(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
      hipGraph_t g;
      HIPOK(hipStreamBeginCapture(cs, hipStreamCaptureModeRelaxed));
      st = launch_fp32(B, o, Ls, cs);
      hipError_t ce = hipStreamEndCapture(cs, &g);
      (void)hipStreamDestroy(cs);
      if (st) return st;
      if (ce != hipSuccess) return fail(KMC_ERR_DEVICE, "graph capture: %s", hipGetErrorString(ce));
      hipError_t ie = hipGraphInstantiate(&B->graph, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (ie != hipSuccess) return fail(KMC_ERR_DEVICE, "graph instantiate: %s", hipGetErrorString(ie));
      B->graph_opts = o;
      B->graph_gen = E.reg_gen;
    }
    HIPOK(hipEventRecord(B->ev0, s));
    HIPOK(hipGraphLaunch(B->graph, s));
    HIPOK(hipEventRecord(B->ev1, s));
  } else {
    HIPOK(hipEventRecord(B->ev0, s));
    st = launch_fp32(B, o, Ls, s);
    if (st) return st;
    HIPOK(hipEventRecord(B->ev1, s));
  }
  B->last_rec = (o.record & 1) != 0;
  B->last_traffic = (o.record & 2) != 0;
  return KMC_OK;
}

}  // namespace
}  // namespace kmc

using namespace kmc;

extern "C" {

const char* kmc_last_error(void) { return g_err.c_str(); }

int kmc_version(void) { return 1; }

int kmc_init(int device) {
  Engine& E = eng();
  std::lock_guard<std::mutex> lk(E.mu);
  if (E.ready) {
    if (device >= 0 && device != E.device)
      return fail(KMC_ERR_ARGS, "engine already bound to device %d", E.device);
    return KMC_OK;
  }
  if (device >= 0) HIPOK(hipSetDevice(device));
  return ensure_init();
}

int kmc_device_info(char* name, int name_len, int* cu_count, int* lds_bytes) {
  Engine& E = eng();
  std::lock_guard<std::mutex> lk(E.mu);
  int st = ensure_init();
  if (st) return st;
  if (name && name_len > 0) snprintf(name, (size_t)name_len, "%s", E.name);
  if (cu_count) *cu_count = E.cu_count;
  if (lds_bytes) *lds_bytes = E.lds_bytes;
  return KMC_OK;
}

int kmc_reset(void) {
  Engine& E = eng();
  std::lock_guard<std::mutex> lk(E.mu);
  if (E.ready) (void)hipStreamSynchronize(E.stream);
  for (Batch* b : E.batches) delete b;
  E.batches.clear();
  for (Layout* l : E.layouts) delete l;
  E.layouts.clear();
  E.esets_host.clear();
  E.eset_layout.clear();
  E.layouts_dirty = E.esets_dirty = true;
  return KMC_OK;
}

int kmc_layout_create(int64_t N, int64_t P, double nu, double kT, double I_0, double R,
                      const double* distances, const double* transitions_constant,
                      double prune_threshold, int32_t* layout_out) {
  Engine& E = eng();
  std::lock_guard<std::mutex> lk(E.mu);
  int st = ensure_init();
  if (st) return st;
  return layout_create_locked(N, P, nu, kT, I_0, R, distances, transitions_constant,
                              prune_threshold, layout_out);
}

int kmc_layout_set_positions(int32_t layout, double ab, const double* positions) {
  Engine& E = eng();
  std::lock_guard<std::mutex> lk(E.mu);
  if (layout < 0 || layout >= (int)E.layouts.size() || !E.layouts[layout])
    return fail(KMC_ERR_ARGS, "bad layout id %d", layout);
  if (!positions || !(ab > 0)) return fail(KMC_ERR_ARGS, "need positions and ab > 0");
  return layout_set_positions_locked(E.layouts[layout], ab, positions);
}

int kmc_batch_update(int32_t batch, const kmc_traj* trajs, int64_t ntraj) {
  Engine& E = eng();
  std::lock_guard<std::mutex> lk(E.mu);
  Batch* B = get_batch(batch);
  if (!B) return fail(KMC_ERR_ARGS, "bad batch id %d", batch);
  return batch_update_locked(B, trajs, ntraj);
}

int kmc_batch_set_state(int32_t batch, const int64_t* h_begin, const int64_t* h_end,
                        const double* time, const int64_t* eo, const uint64_t* occ) {
  Engine& E = eng();
  std::lock_guard<std::mutex> lk(E.mu);
  Batch* B = get_batch(batch);
  if (!B) return fail(KMC_ERR_ARGS, "bad batch id %d", batch);
  return batch_set_state_locked(B, h_begin, h_end, time, eo, occ);
}

int kmc_layout_destroy(int32_t layout) {
  Engine& E = eng();
  std::lock_guard<std::mutex> lk(E.mu);
  return layout_destroy_locked(layout);
}

int kmc_eset_destroy(int32_t eset) {
  Engine& E = eng();
  std::lock_guard<std::mutex> lk(E.mu);
  return eset_destroy_locked(eset);
}

int kmc_eset_create(int32_t layout, int64_t count, const double* E_constant, const double* V,
                    int32_t* first_out) {
  Engine& E = eng();
  std::lock_guard<std::mutex> lk(E.mu);
  int st = ensure_init();
  if (st) return st;
  return eset_create_locked(layout, count, E_constant, V, first_out);
}

int kmc_eset_write(int32_t eset, const double* E_constant, const double* V) {
  Engine& E = eng();
  std::lock_guard<std::mutex> lk(E.mu);
  return eset_write_locked(eset, E_constant, V);
}

int kmc_eset_write_many(int64_t n, const int32_t* ids, const double* E_constant, const double* V) {
  Engine& E = eng();
  std::lock_guard<std::mutex> lk(E.mu);
  int st = ensure_init();
  if (st) return st;
  return eset_write_many_locked(n, ids, E_constant, V);
}

int kmc_batch_set_buckets(int32_t batch, const kmc_bucket* buckets, int32_t nbuckets) {
  Engine& E = eng();
  std::lock_guard<std::mutex> lk(E.mu);
  Batch* B = get_batch(batch);
  if (!B) return fail(KMC_ERR_ARGS, "bad batch id %d", batch);
  if (nbuckets < 0 || nbuckets > KMC_MAX_BUCKETS || (nbuckets > 0 && !buckets))
    return fail(KMC_ERR_ARGS, "nbuckets must be 0..%d", KMC_MAX_BUCKETS);
  for (int b = 0; b < nbuckets; ++b) {
    const kmc_bucket& x = buckets[b];
    if (x.s_max <= 0 || x.s_max > SMAX_FP32)
      return fail(KMC_ERR_ARGS, "bucket %d: s_max must be within 1..%d", b, SMAX_FP32);
    if (x.team != 0 && x.team != 1 && x.team != 2 && x.team != 4)
      return fail(KMC_ERR_ARGS, "bucket %d: team must be 1, 2 or 4", b);
    if (x.min_hops < 0) return fail(KMC_ERR_ARGS, "bucket %d: negative min_hops", b);
    if (x.table_form < 1 || x.table_form > 3)
      return fail(KMC_ERR_ARGS, "bucket %d: table_form must be 1, 2 or 3", b);
    if (x.waves_per_group != 1 && x.waves_per_group != 2 && x.waves_per_group != 4 &&
        x.waves_per_group != 8)
      return fail(KMC_ERR_ARGS, "bucket %d: waves_per_group must be 1, 2, 4 or 8", b);
    if (x.grid < 0) return fail(KMC_ERR_ARGS, "bucket %d: negative grid", b);
    if (x.s_min < 0 || x.s_min > x.s_max)
      return fail(KMC_ERR_ARGS, "bucket %d: s_min must be within 0..s_max", b);
  }
  (void)hipStreamSynchronize(E.stream);
  B->plan.assign(buckets, buckets + nbuckets);
  B->subs_built = false;
  if (B->graph) {
    (void)hipGraphExecDestroy(B->graph);
    B->graph = nullptr;
  }
  return KMC_OK;
}

int kmc_batch_create(const kmc_traj* trajs, int64_t ntraj, int32_t* batch_out) {
  Engine& E = eng();
  std::lock_guard<std::mutex> lk(E.mu);
  int st = ensure_init();
  if (st) return st;
  return batch_create_locked(trajs, ntraj, batch_out);
}

int kmc_batch_run(int32_t batch, const kmc_run_opts* opts) {
  Engine& E = eng();
  std::lock_guard<std::mutex> lk(E.mu);
  int st = ensure_init();
  if (st) return st;
  return batch_run_locked(batch, opts);
}

int kmc_batch_sync(int32_t batch, double* elapsed_ms) {
  Engine& E = eng();
  Batch* B;
  {
    std::lock_guard<std::mutex> lk(E.mu);
    B = get_batch(batch);
    if (!B) return fail(KMC_ERR_ARGS, "bad batch id %d", batch);
  }
  HIPOK(hipEventSynchronize(B->ev1));
  if (elapsed_ms) {
    float ms = 0.f;
    HIPOK(hipEventElapsedTime(&ms, B->ev0, B->ev1));
    *elapsed_ms = ms;
  }
  return KMC_OK;
}

int kmc_batch_wait(int32_t batch, double timeout_ms, double* elapsed_ms) {
  Engine& E = eng();
  Batch* B;
  {
    std::lock_guard<std::mutex> lk(E.mu);
    B = get_batch(batch);
    if (!B) return fail(KMC_ERR_ARGS, "bad batch id %d", batch);
  }
  if (!(timeout_ms >= 0.0)) return fail(KMC_ERR_ARGS, "timeout must be >= 0 ms");
  const auto t0 = std::chrono::steady_clock::now();
  for (;;) {
    const hipError_t q = hipEventQuery(B->ev1);
    if (q == hipSuccess) break;
    if (q != hipErrorNotReady) return fail(KMC_ERR_DEVICE, "event query: %s", hipGetErrorString(q));
    const double ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (ms >= timeout_ms)
      return fail(KMC_ERR_TIMEOUT, "batch %d still running after %.3f ms", batch, ms);
    std::this_thread::sleep_for(std::chrono::microseconds(ms < 2.0 ? 20 : 200));
  }
  if (elapsed_ms) {
    float ms = 0.f;
    HIPOK(hipEventElapsedTime(&ms, B->ev0, B->ev1));
    *elapsed_ms = ms;
  }
  return KMC_OK;
}

int kmc_batch_results(int32_t batch, double* time, int64_t* eo, uint64_t* occ, int32_t* status,
                      int64_t* hops_done, double* occ_time) {
  Engine& E = eng();
  std::lock_guard<std::mutex> lk(E.mu);
  Batch* B = get_batch(batch);
  if (!B) return fail(KMC_ERR_ARGS, "bad batch id %d", batch);
  HIPOK(hipStreamSynchronize(E.stream));
  std::vector<TrajOut> out(B->ntraj);
  HIPOK(hipMemcpy(out.data(), B->out.p, sizeof(TrajOut) * B->ntraj, hipMemcpyDeviceToHost));
  std::vector<double> ot;
  if (occ_time) {
    if (!B->last_rec || B->occ_time.cap < sizeof(double) * NMAX * (size_t)B->ntraj)
      return fail(KMC_ERR_ARGS, "the last run of this batch did not record occupation time (record bit 0)");
    ot.resize((size_t)NMAX * B->ntraj);
    HIPOK(hipMemcpy(ot.data(), B->occ_time.p, sizeof(double) * ot.size(), hipMemcpyDeviceToHost));
  }
  for (int64_t i = 0; i < B->ntraj; ++i) {
    const int64_t j = B->perm[i];   // caller's index
    const TrajOut& r = out[i];
    const int P = eng().layouts[B->host[i].layout]->P;
    if (time) time[j] = r.time;
    if (status) status[j] = r.status;
    if (hops_done) hops_done[j] = r.hops_done;
    if (occ)
      for (int w = 0; w < OCCW; ++w) occ[j * OCCW + w] = r.occ[w];
    if (eo)
      for (int e = 0; e < PMAX; ++e) eo[j * PMAX + e] = e < P ? r.eo[e] : 0;
    if (occ_time) memcpy(occ_time + (size_t)j * NMAX, ot.data() + (size_t)i * NMAX, sizeof(double) * NMAX);
  }
  return KMC_OK;
}

int kmc_batch_timing(int32_t batch, double* seconds) {
  Engine& E = eng();
  std::lock_guard<std::mutex> lk(E.mu);
  Batch* B = get_batch(batch);
  if (!B) return fail(KMC_ERR_ARGS, "bad batch id %d", batch);
  if (!seconds) return fail(KMC_ERR_ARGS, "null output");
  int khz = 0;
  HIPOK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, E.device));
  if (khz <= 0) return fail(KMC_ERR_DEVICE, "no wall-clock rate");
  HIPOK(hipStreamSynchronize(E.stream));
  std::vector<TrajOut> out(B->ntraj);
  HIPOK(hipMemcpy(out.data(), B->out.p, sizeof(TrajOut) * B->ntraj, hipMemcpyDeviceToHost));
  for (int64_t i = 0; i < B->ntraj; ++i) seconds[B->perm[i]] = out[i].ticks / (khz * 1e3);
  return KMC_OK;
}

int kmc_batch_traffic(int32_t batch, double* traffic, int64_t* stride_out) {
  Engine& E = eng();
  std::lock_guard<std::mutex> lk(E.mu);
  Batch* B = get_batch(batch);
  if (!B) return fail(KMC_ERR_ARGS, "bad batch id %d", batch);
  if (!B->last_traffic)
    return fail(KMC_ERR_ARGS, "the last run of this batch recorded no traffic (FP32 run with record bit 1)");
  const int64_t stride = (int64_t)B->smax * B->smax;
  if (stride_out) *stride_out = stride;
  if (!traffic) return KMC_OK;
  HIPOK(hipStreamSynchronize(E.stream));
  std::vector<double> t((size_t)stride * B->ntraj);
  HIPOK(hipMemcpy(t.data(), B->traffic.p, sizeof(double) * t.size(), hipMemcpyDeviceToHost));
  for (int64_t i = 0; i < B->ntraj; ++i)
    memcpy(traffic + (size_t)B->perm[i] * stride, t.data() + (size_t)i * stride,
           sizeof(double) * stride);
  return KMC_OK;
}

int kmc_probe_fp64_phases(int32_t batch, uint64_t* cycles, int32_t n) {
  Engine& E = eng();
  std::lock_guard<std::mutex> lk(E.mu);
  int st = ensure_init();
  if (st) return st;
  Batch* B = get_batch(batch);
  if (!B) return fail(KMC_ERR_ARGS, "bad batch id %d", batch);
  if (!cycles || n < PLAIN_PHASES) return fail(KMC_ERR_ARGS, "need %d phase slots", PLAIN_PHASES);
  if ((st = sync_registry())) return st;
  DevBuf ph;
  if ((st = ph.alloc(sizeof(uint64_t) * PLAIN_PHASES))) return st;
  HIPOK(hipMemset(ph.p, 0, sizeof(uint64_t) * PLAIN_PHASES));
  kmc_run_opts o{};
  o.precision = KMC_FP64_DENSE;
  if ((st = enqueue_fp64(B, o, E.stream, nullptr, nullptr, nullptr, nullptr, nullptr,
                         ph.as<uint64_t>())))
    return st;
  HIPOK(hipStreamSynchronize(E.stream));
  HIPOK(hipMemcpy(cycles, ph.p, sizeof(uint64_t) * PLAIN_PHASES, hipMemcpyDeviceToHost));
  B->last_rec = B->last_traffic = false;
  return KMC_OK;
}

int kmc_probe_fp32_phases(int32_t batch, const kmc_run_opts* opts, uint64_t* cycles, int32_t n) {
  Engine& E = eng();
  std::lock_guard<std::mutex> lk(E.mu);
  int st = ensure_init();
  if (st) return st;
  Batch* B = get_batch(batch);
  if (!B) return fail(KMC_ERR_ARGS, "bad batch id %d", batch);
  if (!cycles || n < FP32_PHASES) return fail(KMC_ERR_ARGS, "need %d phase slots", FP32_PHASES);
  kmc_run_opts o{};
  if (opts) o = *opts;
  if (o.precision != KMC_FP32_BATCHED || o.record || o.use_graph || o.resume)
    return fail(KMC_ERR_UNSUPPORTED, "phase timing: FP32 batched runs without records, graph or resume");
  if ((st = sync_registry())) return st;
  DevBuf ph;
  if ((st = ph.alloc(sizeof(uint64_t) * FP32_PHASES))) return st;
  HIPOK(hipMemset(ph.p, 0, sizeof(uint64_t) * FP32_PHASES));
  std::vector<Fp32Launch> Ls;
  if ((st = prepare_fp32(B, o, Ls, ph.as<uint64_t>()))) return st;
  if (Ls.size() > 1 && (st = ensure_side_streams(B, (int)Ls.size()))) return st;
  if ((st = launch_fp32(B, o, Ls, E.stream))) return st;
  HIPOK(hipStreamSynchronize(E.stream));
  HIPOK(hipMemcpy(cycles, ph.p, sizeof(uint64_t) * FP32_PHASES, hipMemcpyDeviceToHost));
  B->last_rec = B->last_traffic = false;
  return KMC_OK;
}

int kmc_batch_destroy(int32_t batch) {
  Engine& E = eng();
  std::lock_guard<std::mutex> lk(E.mu);
  Batch* B = get_batch(batch);
  if (!B) return fail(KMC_ERR_ARGS, "bad batch id %d", batch);
  (void)hipStreamSynchronize(E.stream);
  delete B;
  E.batches[batch] = nullptr;
  return KMC_OK;
}

}  // extern "C" (reopened below)

namespace {

// One wave per network; n networks of layouts[i] / esets[i].  avg_current: the
// calcAverageCurrent evaluation (currents into eo_out), else probSimulate for
// `steps` steps (occupation in/out).
int meanfield_locked(bool avg_current, int32_t n, const int32_t* layouts, const int32_t* esets,
                     int64_t steps, double* occupation, double* eo_out, double* se_out,
                     double* time_out, double* avg_out, double* traffic_out) {
  Engine& E = eng();
  int st = ensure_init();
  if (st) return st;
  if (n < 0 || (n > 0 && (!layouts || !esets || !occupation || !eo_out)))
    return fail(KMC_ERR_ARGS, "null argument");
  if (steps < 0) return fail(KMC_ERR_ARGS, "steps must be >= 0");
  if (n == 0) return KMC_OK;
  int smax = 0;
  for (int32_t i = 0; i < n; ++i) {
    const int32_t l = layouts[i], e = esets[i];
    if (l < 0 || l >= (int)E.layouts.size() || !E.layouts[l])
      return fail(KMC_ERR_ARGS, "network %d: no layout %d", i, l);
    if (e < 0 || e >= (int)E.eset_layout.size() || E.eset_layout[e] != l)
      return fail(KMC_ERR_ARGS, "network %d: energy set %d does not belong to layout %d", i, e, l);
    smax = std::max(smax, E.layouts[l]->S);
  }
  if ((st = sync_registry())) return st;
  const int C = (smax + 63) / 64;
  const size_t na = (size_t)n * NMAX, ne = (size_t)n * PMAX;
  const int64_t tstride = (int64_t)smax * smax;
  DevBuf ids, occ, eo, se, tm, avg, tr;
  std::vector<int32_t> both((size_t)2 * n);
  std::copy(layouts, layouts + n, both.begin());
  std::copy(esets, esets + n, both.begin() + n);
  if ((st = upload(ids, both)) || (st = occ.alloc(na * sizeof(double))) ||
      (st = eo.alloc(ne * sizeof(double))) || (st = tm.alloc((size_t)n * sizeof(double))) ||
      (se_out && (st = se.alloc(na * sizeof(double)))) ||
      (avg_out && (st = avg.alloc(na * sizeof(double)))) ||
      (traffic_out && (st = tr.alloc((size_t)n * tstride * sizeof(double)))))
    return st;
  HIPOK(hipMemcpyAsync(occ.p, occupation, na * sizeof(double), hipMemcpyHostToDevice, E.stream));
  HIPOK(hipMemsetAsync(eo.p, 0, ne * sizeof(double), E.stream));
  if (traffic_out) HIPOK(hipMemsetAsync(tr.p, 0, (size_t)n * tstride * sizeof(double), E.stream));
  MeanfieldParams mp{};
  mp.layouts = E.layouts_dev.as<LayoutDev>();
  mp.esets = E.esets_dev.as<EsetDev>();
  mp.layout_of = ids.as<int32_t>();
  mp.eset_of = ids.as<int32_t>() + n;
  mp.occ = occ.as<double>();
  mp.eo = eo.as<double>();
  mp.se = se_out ? se.as<double>() : nullptr;
  mp.time = tm.as<double>();
  mp.avg = avg_out ? avg.as<double>() : nullptr;
  mp.traffic = traffic_out ? tr.as<double>() : nullptr;
  mp.tstride = tstride;
  mp.steps = steps;
  mp.n = n;
  using K = void (*)(MeanfieldParams);
  K k = nullptr;
  if (avg_current)
    k = C == 1 ? (K)kmc_avg_current_kernel<1> : C == 2 ? (K)kmc_avg_current_kernel<2> : (K)kmc_avg_current_kernel<3>;
  else
    k = C == 1 ? (K)kmc_meanfield_kernel<1> : C == 2 ? (K)kmc_meanfield_kernel<2> : (K)kmc_meanfield_kernel<3>;
  hipLaunchKernelGGL(k, dim3(n), dim3(64), 0, E.stream, mp);
  HIPOK(hipGetLastError());
  if (!avg_current) {
    HIPOK(hipMemcpyAsync(occupation, occ.p, na * sizeof(double), hipMemcpyDeviceToHost, E.stream));
    if (time_out) HIPOK(hipMemcpyAsync(time_out, tm.p, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, E.stream));
    if (se_out) HIPOK(hipMemcpyAsync(se_out, se.p, na * sizeof(double), hipMemcpyDeviceToHost, E.stream));
    if (avg_out) HIPOK(hipMemcpyAsync(avg_out, avg.p, na * sizeof(double), hipMemcpyDeviceToHost, E.stream));
    if (traffic_out)
      HIPOK(hipMemcpyAsync(traffic_out, tr.p, (size_t)n * tstride * sizeof(double), hipMemcpyDeviceToHost, E.stream));
  }
  HIPOK(hipMemcpyAsync(eo_out, eo.p, ne * sizeof(double), hipMemcpyDeviceToHost, E.stream));
  HIPOK(hipStreamSynchronize(E.stream));
  return KMC_OK;
}

}  // namespace

extern "C" {

int kmc_meanfield_simulate(int32_t n, const int32_t* layouts, const int32_t* esets, int64_t steps,
                           double* occupation, double* electrode_occupation, double* site_energies,
                           double* time, double* average_occupation, double* traffic) {
  Engine& E = eng();
  std::lock_guard<std::mutex> lk(E.mu);
  return meanfield_locked(false, n, layouts, esets, steps, occupation, electrode_occupation,
                          site_energies, time, average_occupation, traffic);
}

int kmc_average_current(int32_t n, const int32_t* layouts, const int32_t* esets,
                        const double* occupation, double* currents) {
  Engine& E = eng();
  std::lock_guard<std::mutex> lk(E.mu);
  return meanfield_locked(true, n, layouts, esets, 0, const_cast<double*>(occupation), currents,
                          nullptr, nullptr, nullptr, nullptr);
}

int kmc_simulate_discrete(int64_t N, int64_t P, double nu, double kT, double I_0, double R,
                          const double* distances, const double* E_constant,
                          double* site_energies, const double* transitions_constant,
                          uint8_t* occupation, int64_t* electrode_occupation, int64_t hops,
                          int record, double* traffic, double* occupations_in_time,
                          const double* uniforms, uint64_t seed, uint64_t hop0,
                          int32_t* events_out, double* totals_out, double* time_out) {
  Engine& E = eng();
  std::lock_guard<std::mutex> lk(E.mu);
  int st = ensure_init();
  if (st) return st;
  if (!site_energies || !occupation || (P > 0 && !electrode_occupation) || !E_constant)
    return fail(KMC_ERR_ARGS, "null argument");
  if (hops < 0 || hops > INT32_MAX) return fail(KMC_ERR_ARGS, "hops out of range");
  if (N + P > SMAX_FP64)
    return fail(KMC_ERR_LIMIT, "FP64 dense kernel supports S <= %d (got %lld)", SMAX_FP64, (long long)(N + P));
  int32_t lid, eid, bid;
  if ((st = layout_create_locked(N, P, nu, kT, I_0, R, distances, transitions_constant, 0.0, &lid)))
    return st;
  // The layout/energy set/batch are private to this call and released below.
  auto cleanup = [&](int code) {
    (void)hipStreamSynchronize(E.stream);
    if (bid >= 0 && bid < (int)E.batches.size()) {
      delete E.batches[bid];
      E.batches[bid] = nullptr;
    }
    layout_destroy_locked(lid);
    return code;
  };
  eid = -1;
  bid = -1;
  if ((st = eset_create_locked(lid, 1, E_constant, site_energies + N, &eid))) return cleanup(st);
  kmc_traj tr{};
  tr.layout = lid;
  tr.eset = eid;
  tr.seed = seed;
  tr.hop0 = hop0;
  tr.prehops = 0;
  tr.hops = hops;
  for (int64_t i = 0; i < N; ++i)
    if (occupation[i]) tr.occ[i >> 6] |= 1ull << (i & 63);
  int32_t b;
  if ((st = batch_create_locked(&tr, 1, &b))) return cleanup(st);
  bid = b;
  if ((st = sync_registry())) return cleanup(st);
  Batch* B = E.batches[bid];
  DevBuf du, dev_ev, dtot, dtr, dse;
  const int S = (int)(N + P);
  if (uniforms) {
    if ((st = du.alloc(sizeof(double) * 2 * hops))) return cleanup(st);
    if (hops) HIPOK(hipMemcpy(du.p, uniforms, sizeof(double) * 2 * hops, hipMemcpyHostToDevice));
  }
  if (events_out && (st = dev_ev.alloc(sizeof(int32_t) * hops))) return cleanup(st);
  if (totals_out && (st = dtot.alloc(sizeof(double) * hops))) return cleanup(st);
  if (record && traffic) {
    if ((st = dtr.alloc(sizeof(double) * S * S))) return cleanup(st);
    HIPOK(hipMemset(dtr.p, 0, sizeof(double) * S * S));
  }
  if ((st = dse.alloc(sizeof(double) * NMAX))) return cleanup(st);
  kmc_run_opts o{};
  o.precision = KMC_FP64_DENSE;
  o.record = record ? 1 : 0;
  o.post_hop_record = 1;
  if ((st = enqueue_fp64(B, o, E.stream, uniforms ? du.as<double>() : nullptr,
                         events_out ? dev_ev.as<int32_t>() : nullptr,
                         totals_out ? dtot.as<double>() : nullptr,
                         (record && traffic) ? dtr.as<double>() : nullptr, dse.as<double>())))
    return cleanup(st);
  hipError_t se = hipStreamSynchronize(E.stream);
  if (se != hipSuccess) return cleanup(fail(KMC_ERR_DEVICE, "kernel: %s", hipGetErrorString(se)));
  TrajOut r;
  HIPOK(hipMemcpy(&r, B->out.p, sizeof r, hipMemcpyDeviceToHost));
  for (int64_t i = 0; i < N; ++i) occupation[i] = (r.occ[i >> 6] >> (i & 63)) & 1ull;
  for (int64_t e = 0; e < P; ++e) electrode_occupation[e] += r.eo[e];
  if (time_out) *time_out = r.time;
  std::vector<double> sev(NMAX);
  HIPOK(hipMemcpy(sev.data(), dse.p, sizeof(double) * NMAX, hipMemcpyDeviceToHost));
  if (hops > 0)
    for (int64_t i = 0; i < N; ++i) site_energies[i] = sev[i];
  if (events_out && hops) HIPOK(hipMemcpy(events_out, dev_ev.p, sizeof(int32_t) * hops, hipMemcpyDeviceToHost));
  if (totals_out && hops) HIPOK(hipMemcpy(totals_out, dtot.p, sizeof(double) * hops, hipMemcpyDeviceToHost));
  if (record) {
    if (traffic) HIPOK(hipMemcpy(traffic, dtr.p, sizeof(double) * S * S, hipMemcpyDeviceToHost));
    if (occupations_in_time) {
      std::vector<double> ot(NMAX);
      HIPOK(hipMemcpy(ot.data(), B->occ_time.p, sizeof(double) * NMAX, hipMemcpyDeviceToHost));
      for (int64_t i = 0; i < N; ++i) occupations_in_time[i] = ot[i];
    }
  }
  const int status = r.status;
  cleanup(KMC_OK);
  if (status != KMC_OK)
    return fail(status, "total rate became zero after %lld hops (reference: ZeroDivisionError)",
                (long long)r.hops_done);
  return KMC_OK;
}

// kmc_simulate_discrete without allocation: the same kernel and results, with
// every device buffer in the `scratch` array (its address from set-up) and the
// host side working on the caller's arrays and stack records only.
int kmc_simulate_discrete_noalloc(int64_t N, int64_t P, double nu, double kT, double I_0,
                                  double R, const double* distances, const double* E_constant,
                                  double* site_energies, const double* transitions_constant,
                                  uint8_t* occupation, int64_t* electrode_occupation,
                                  int64_t hops, int record, double* traffic,
                                  double* occupations_in_time, const double* uniforms,
                                  uint64_t seed, uint64_t hop0, int32_t* events_out,
                                  double* totals_out, double* time_out) {
  Engine& E = eng();
  std::lock_guard<std::mutex> lk(E.mu);
  int st = ensure_init();
  if (st) return st;
  char* base = static_cast<char*>(noalloc_scratch[E.device]);
  if (!base) return fail(KMC_ERR_DEVICE, "scratch array not set up on device %d", E.device);
  if (!site_energies || !occupation || (P > 0 && !electrode_occupation) || !E_constant)
    return fail(KMC_ERR_ARGS, "null argument");
  if ((st = check_network(N, P, KMC_NOALLOC_MAX_SITES, nu, kT, I_0, R, distances,
                          transitions_constant)))
    return st;
  if (hops < 0 || hops > INT32_MAX) return fail(KMC_ERR_ARGS, "hops out of range");
  if ((uniforms || events_out || totals_out) && hops > KMC_NOALLOC_MAX_LOGGED_HOPS)
    return fail(KMC_ERR_LIMIT, "replayed uniforms or event logs hold at most %d hops here (got %lld)",
                KMC_NOALLOC_MAX_LOGGED_HOPS, (long long)hops);
  for (int64_t i = 0; i < N; ++i)
    if (!std::isfinite(E_constant[i])) return fail(KMC_ERR_ARGS, "E_constant[%lld] is not finite", (long long)i);
  for (int64_t e = 0; e < P; ++e)
    if (!std::isfinite(site_energies[N + e]))
      return fail(KMC_ERR_ARGS, "V[%lld] is not finite", (long long)e);
  const int S = (int)(N + P), n = (int)N;
  const size_t SS = (size_t)S * S;
  hipStream_t s = E.stream;
  LayoutDev* dL = reinterpret_cast<LayoutDev*>(base + NA_DESC);
  EsetDev* dE = reinterpret_cast<EsetDev*>(base + NA_DESC + 256);
  TrajDev* dT = reinterpret_cast<TrajDev*>(base + NA_DESC + 512);
  TrajOut* dO = reinterpret_cast<TrajOut*>(base + NA_DESC + 640);
  double* dEc = reinterpret_cast<double*>(base + NA_E);
  double* dV = reinterpret_cast<double*>(base + NA_V);
  double* dd = reinterpret_cast<double*>(base + NA_D);
  double* dtc = reinterpret_cast<double*>(base + NA_TC);
  double* dird = reinterpret_cast<double*>(base + NA_IRD);
  double* dinvd = reinterpret_cast<double*>(base + NA_INVD);
  double* dtr = reinterpret_cast<double*>(base + NA_TRAFFIC);
  double* dse = reinterpret_cast<double*>(base + NA_SE);
  double* dot = reinterpret_cast<double*>(base + NA_OT);
  double* du = reinterpret_cast<double*>(base + NA_U);
  int32_t* dev = reinterpret_cast<int32_t*>(base + NA_EV);
  double* dtot = reinterpret_cast<double*>(base + NA_TOT);
  static_assert(sizeof(LayoutDev) <= 256 && sizeof(EsetDev) <= 256 && sizeof(TrajDev) <= 128 &&
                    640 + sizeof(TrajOut) <= 1024,
                "descriptor slots");
  LayoutDev Ld{};
  Ld.N = n;
  Ld.P = (int)P;
  Ld.S = S;
  Ld.nu = nu;
  Ld.kT = kT;
  Ld.I0 = I_0;
  Ld.R = R;
  Ld.I0R = I_0 * R;
  Ld.tc64 = dtc;
  Ld.ird64 = dird;
  Ld.invd64 = dinvd;
  Ld.dist64 = dd;
  const EsetDev Ed{dEc, dV};
  TrajDev Td{};
  Td.layout = 0;
  Td.eset = 0;
  Td.seed = seed;
  Td.hop0 = hop0;
  Td.hops = hops;
  Td.h_end = hops;
  for (int64_t i = 0; i < N; ++i)
    if (occupation[i]) Td.occ[i >> 6] |= 1ull << (i & 63);
  const bool rec = record != 0;
  HIPOK(hipMemcpyAsync(dL, &Ld, sizeof Ld, hipMemcpyHostToDevice, s));
  HIPOK(hipMemcpyAsync(dE, &Ed, sizeof Ed, hipMemcpyHostToDevice, s));
  HIPOK(hipMemcpyAsync(dT, &Td, sizeof Td, hipMemcpyHostToDevice, s));
  if (N) HIPOK(hipMemcpyAsync(dEc, E_constant, sizeof(double) * N, hipMemcpyHostToDevice, s));
  if (P) HIPOK(hipMemcpyAsync(dV, site_energies + N, sizeof(double) * P, hipMemcpyHostToDevice, s));
  HIPOK(hipMemcpyAsync(dd, distances, sizeof(double) * SS, hipMemcpyHostToDevice, s));
  HIPOK(hipMemcpyAsync(dtc, transitions_constant, sizeof(double) * SS, hipMemcpyHostToDevice, s));
  if (uniforms && hops)
    HIPOK(hipMemcpyAsync(du, uniforms, sizeof(double) * 2 * hops, hipMemcpyHostToDevice, s));
  {
    double I0R = I_0 * R;
    int clear = rec ? 1 : 0, nn = n, SSi = S;
    void* a[] = {&dd, &dird, &dinvd, &dtr, &dot, &nn, &SSi, &I0R, &clear};
    HIPOK(hipLaunchKernel((const void*)kmc_noalloc_prep_kernel, dim3((unsigned)((SS + 255) / 256)),
                          dim3(256), a, 0, s));
  }
  PlainParams p{};
  p.trajs = dT;
  p.layouts = dL;
  p.esets = dE;
  p.out = dO;
  p.uniforms = uniforms ? du : nullptr;
  p.events = events_out ? dev : nullptr;
  p.totals = totals_out ? dtot : nullptr;
  p.traffic = (rec && traffic) ? dtr : nullptr;
  p.occ_time = rec ? dot : nullptr;
  p.se_out = dse;
  if ((st = launch_plain(p, S, n, rec, 1, s))) return st;
  TrajOut r;
  HIPOK(hipMemcpyAsync(&r, dO, sizeof r, hipMemcpyDeviceToHost, s));
  if (hops > 0 && N) HIPOK(hipMemcpyAsync(site_energies, dse, sizeof(double) * N, hipMemcpyDeviceToHost, s));
  if (events_out && hops)
    HIPOK(hipMemcpyAsync(events_out, dev, sizeof(int32_t) * hops, hipMemcpyDeviceToHost, s));
  if (totals_out && hops)
    HIPOK(hipMemcpyAsync(totals_out, dtot, sizeof(double) * hops, hipMemcpyDeviceToHost, s));
  if (rec && traffic) HIPOK(hipMemcpyAsync(traffic, dtr, sizeof(double) * SS, hipMemcpyDeviceToHost, s));
  if (rec && occupations_in_time && N)
    HIPOK(hipMemcpyAsync(occupations_in_time, dot, sizeof(double) * N, hipMemcpyDeviceToHost, s));
  hipError_t se = hipStreamSynchronize(s);
  if (se != hipSuccess) return fail(KMC_ERR_DEVICE, "kernel: %s", hipGetErrorString(se));
  for (int64_t i = 0; i < N; ++i) occupation[i] = (r.occ[i >> 6] >> (i & 63)) & 1ull;
  for (int64_t e = 0; e < P; ++e) electrode_occupation[e] += r.eo[e];
  if (time_out) *time_out = r.time;
  if (r.status != KMC_OK)
    return fail(r.status, "total rate became zero after %lld hops (reference: ZeroDivisionError)",
                (long long)r.hops_done);
  return KMC_OK;
}

}  // extern "C"
