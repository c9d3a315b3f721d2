// C ABI and host orchestration of one batched evaluation (include/xreac.h).
//
// A call evaluates one or more independent systems packed atom-wise; a
// single system is a batch of one.  Pipeline on the context's stream (host
// synchronisation points marked *):
//   system table upload; bounds* (only with nonperiodic directions) -> bin grids
//   bin, scatter -> neighbor count -> scan -> neighbor fill (canonical rows)
//   edge pass (H values, active-bond flags) -> scan -> bond compaction
//   QEq CG (two RHS, per-system convergence)*  -> nonbonded energy/forces
//   bond orders, lone pair, over/under, special terms
//   angles, torsions, hydrogen bonds   (forward + adjoint deposits)
//   bond-order adjoint -> forces -> per-system energy reduction*
// Edge and bond buffers are sized from the previous call's per-atom counts;
// an overflow is detected on the device and the call repeats with sizes read
// back first (two more synchronisations).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/xreac.h"
#include "integrate.h"
#include "launch.h"
#include "profile.h"
#include "qeq.h"
#include "state.h"

static_assert(XREAC_ATOM_PARAMS == AP_COUNT, "atom table layout");
static_assert(XREAC_PAIR_PARAMS == PP_COUNT, "pair table layout");
static_assert(XREAC_COMPONENTS == NCOMP, "component count");
static_assert(XREAC_TIMERS == XT_COUNT, "timer count");

thread_local XrProfiler* xr_active_profiler = nullptr;

struct xreac_ff {
  FFDev dev{};
  std::vector<void*> allocations;
  double cutoff = 0.0;
};

namespace {

// Grow-only device buffer.
struct Buffer {
  void* ptr = nullptr;
  size_t bytes = 0;
  hipError_t ensure(size_t need) {
    if (need <= bytes) return hipSuccess;
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
    const size_t grow = need + need / 4 + 256;
    hipError_t e = hipMalloc(&ptr, grow);
    if (e == hipSuccess) bytes = grow;
    return e;
  }
  ~Buffer() {
    if (ptr) (void)hipFree(ptr);
  }
};

// Sequential carving of one buffer into aligned typed arrays.
struct Carver {
  char* base;
  size_t off = 0;
  template <class T>
  T* take(size_t count) {
    off = (off + 255) & ~size_t(255);
    T* p = reinterpret_cast<T*>(base + off);
    off += count * sizeof(T);
    return p;
  }
};

template <class T>
size_t carve_size(size_t& off, size_t count) {
  off = (off + 255) & ~size_t(255);
  off += count * sizeof(T);
  return off;
}

// Sequential carving of the static scratch array for xreac_evaluate_static:
// each workspace the evaluation needs is taken from the front of what is left.
struct Arena {
  char* base = nullptr;
  size_t cap = 0;
  size_t used = 0;
  size_t wanted = 0;  // bytes the evaluation asked for in total (reported when cap is exceeded)
};

// Workspace of xreac_evaluate_static.  A host reference to this name is a
// host-side stand-in, not a device address: host code uses only the address
// hipGetSymbolAddress returns for the current device (scratch_ptr below).
static __device__ __align__(256) unsigned char scratch[XREAC_STATIC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
void* g_scratch_ptr[kMaxDevices] = {};  // device address of scratch on each device, once resolved

// One evaluation at a time per device, whatever the context: evaluations of
// two contexts whose kernels overlapped on one device gave rare run-to-run
// differences (the serving stream's two force-field groups: a few of 20,000
// requests per run, occasionally a wrong evaluation), while serialized ones
// are bitwise reproducible.  The integrators stay outside the lock.
std::mutex g_eval_lock[kMaxDevices];

// The kept skin rows are geometry only: they depend on the cutoff, the skin
// and the systems' atom ranges and cells (types are compared on the device).
struct SkinKey {
  double rc = 0.0;
  double skin = 0.0;
  std::vector<int32_t> offsets;
  std::vector<double> cells;
  std::vector<int32_t> pbc;
  bool operator==(const SkinKey& o) const {
    return rc == o.rc && skin == o.skin && offsets == o.offsets && cells == o.cells && pbc == o.pbc;
  }
};

// Workspaces and sizing history of one evaluation pipeline: the context's
// own, and one per captured step (whose recorded addresses must stay valid
// while the context evaluates other batches).
struct EvalState {
  Buffer atom_buf, bin_buf, edge_buf, bond_buf, ctl_buf, slab_buf;
  // Skin reuse (xreac_options.neighbor_skin): the kept rows within rc + 2 skin
  // (slab of skin_stride keys per atom), their lengths, and the positions and
  // types they were built at; valid for calls of the same shape as the last.
  Buffer skin_slab_buf, skin_counts_buf, skin_xref_buf, skin_tref_buf;
  Buffer guess_buf;  // the warm-start guess as it was before an attempt that may be repeated
  SkinKey skin_key;
  bool skin_valid = false;
  int skin_stride = 0;
  char* ctl_host = nullptr;  // pinned mirror of the control block
  size_t ctl_host_bytes = 0;
  int last_cg_launches = 0;
  int cg_history[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // launches of the last eight solves (0: none yet)
  int cg_history_next = 0;
  double m_per_atom = 0.0;  // edges and bonds per atom of the last evaluation
  double b_per_atom = 0.0;
  int max_row = 0;  // longest neighbor row of the last built list
  // last evaluation
  int n = 0;
  int64_t m = 0, nb = 0;
  int* row_ptr = nullptr;
  int* nbr_j = nullptr;
  int* nbr_s = nullptr;
  int* bond_ptr = nullptr;
  int* bond_i = nullptr;
  int* bond_j = nullptr;
  int* bond_code = nullptr;
  double* bond_BO = nullptr;
  EvalState() = default;
  EvalState(const EvalState&) = delete;
  EvalState& operator=(const EvalState&) = delete;
  ~EvalState() {
    if (ctl_host) (void)hipHostFree(ctl_host);
  }
  // Sizing history only (what the next evaluation's buffers are sized from).
  void take_hints(const EvalState& o) {
    last_cg_launches = o.last_cg_launches;
    std::copy(o.cg_history, o.cg_history + 8, cg_history);
    cg_history_next = o.cg_history_next;
    m_per_atom = o.m_per_atom;
    b_per_atom = o.b_per_atom;
    max_row = o.max_row;
  }
};

}  // namespace

struct xreac_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string message;
  std::vector<std::string> sys_messages;  // per system of the last batch
  EvalState es;
  int64_t neighbor_list_builds = 0;
  XrProfiler prof;
  // Second stream (default priority) for the charge-independent bonded terms,
  // which run while the QEq solve occupies the context's stream.
  hipStream_t side = nullptr;
  hipEvent_t bonds_ready = nullptr, side_done = nullptr;
  // QEq read-back: pinned slots and their copy events
  int* poll_host = nullptr;
  hipEvent_t poll_ev[2] = {nullptr, nullptr};
  void* scratch = nullptr;  // this device's scratch array (resolved at creation)
};

namespace {

constexpr int kBondScalars = 27;  // per-bond double arrays carved below
constexpr int kAtomScalars = 22;  // per-atom double arrays carved below

int fail(xreac_ctx* c, int code, const std::string& msg) {
  c->message = msg;
  return code;
}

int hip_fail(xreac_ctx* c, hipError_t e, const char* where) {
  c->message = std::string("HIP error in ") + where + ": " + hipGetErrorString(e);
  return e == hipErrorOutOfMemory ? XREAC_NOMEM : XREAC_DEVICE;
}

#define HIPCHK(expr, where)                          \
  do {                                               \
    hipError_t e_ = (expr);                          \
    if (e_ != hipSuccess) return hip_fail(ctx, e_, where); \
  } while (0)

std::string error_text(unsigned bits) {
  if (bits & XR_ERR_POSITIONS) return "positions must be a finite (N, 3) array in Angstrom";
  if (bits & XR_ERR_COINCIDENT) return "Coincident atoms are not supported";
  // a malformed supplied list comes next: it can also trip every later check
  if (bits & XR_ERR_NEIGHBORS)
    return "Neighbor atom indices are out of range, an image shift exceeds 511 cells, or the neighbor row "
           "pointers are malformed";
  if (bits & XR_ERR_SHIFT_RANGE)
    return "Positions lie more than 511 periodic images from the cell; wrap them before evaluating";
  if (bits & XR_ERR_ROW_TOO_LONG) return "More than 2048 neighbors within the cutoff for one atom";
  if (bits & XR_ERR_BOC3) return "Invalid p_boc3 mixing for an active bond-order correction";
  if (bits & XR_ERR_BOC4) return "Invalid p_boc4 mixing for an active bond-order correction";
  if (bits & XR_ERR_BOC5) return "Invalid p_boc5 mixing for an active bond-order correction";
  if (bits & XR_ERR_COLLINEAR)
    return "Collinear atoms in an active torsion: the dihedral derivative is undefined; perturb the geometry";
  if (bits & XR_ERR_REVERSE) return "Internal error: bond without a reverse edge";
  return "Non-finite energy, charges, or forces; check the geometry";
}

void invert3(const double* a, double* inv) {
  const double det = a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) +
                     a[2] * (a[3] * a[7] - a[4] * a[6]);
  inv[0] = (a[4] * a[8] - a[5] * a[7]) / det;
  inv[1] = (a[2] * a[7] - a[1] * a[8]) / det;
  inv[2] = (a[1] * a[5] - a[2] * a[4]) / det;
  inv[3] = (a[5] * a[6] - a[3] * a[8]) / det;
  inv[4] = (a[0] * a[8] - a[2] * a[6]) / det;
  inv[5] = (a[2] * a[3] - a[0] * a[5]) / det;
  inv[6] = (a[3] * a[7] - a[4] * a[6]) / det;
  inv[7] = (a[1] * a[6] - a[0] * a[7]) / det;
  inv[8] = (a[0] * a[4] - a[1] * a[3]) / det;
}

// Zero up to eight byte ranges (multiples of 4 bytes) in one launch.
struct ZeroRanges {
  unsigned* ptr[8];
  long long words[8];
  int count;
};

__global__ void k_zero_ranges(ZeroRanges z) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (int r = 0; r < z.count; ++r) {
    unsigned* p = z.ptr[r];
    for (long long w = blockIdx.x * (long long)blockDim.x + threadIdx.x; w < z.words[r]; w += stride) p[w] = 0u;
  }
}

// Copy up to eight byte ranges (multiples of 4 bytes) in one launch.
struct CopyRanges {
  const unsigned* src[8];
  unsigned* dst[8];
  long long words[8];
  int count;
};

__global__ void k_copy_ranges(CopyRanges c) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (int r = 0; r < c.count; ++r)
    for (long long w = blockIdx.x * (long long)blockDim.x + threadIdx.x; w < c.words[r]; w += stride)
      c.dst[r][w] = c.src[r][w];
}

// Record total = ptr[n] and clamp every pointer to `cap`, so later kernels
// stay inside buffers sized for `cap` entries even when the total exceeds it.
template <class T>
hipError_t upload(xreac_ff* ff, const T* src, size_t count, const T** dst) {
  if (count == 0) count = 1;
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, count * sizeof(T));
  if (e != hipSuccess) return e;
  ff->allocations.push_back(p);
  if (src) {
    e = hipMemcpy(p, src, count * sizeof(T), hipMemcpyHostToDevice);
  } else {
    e = hipMemset(p, 0, count * sizeof(T));
  }
  *dst = static_cast<const T*>(p);
  return e;
}

}  // namespace

extern "C" {

int xreac_ff_create(const xreac_ff_tables* t, xreac_ff** out) {
  if (!t || !out || t->ntypes <= 0 || !t->general || !t->atom || !t->pair || !t->angle_offsets) return XREAC_INVALID;
  auto* ff = new xreac_ff();
  const size_t nt = (size_t)t->ntypes;
  ff->dev.nt = t->ntypes;
  ff->dev.vdw_type = t->vdw_type;
  ff->cutoff = t->general[12];
  hipError_t e = hipSuccess;
  const size_t nang = (size_t)t->angle_offsets[nt * nt * nt];
  if (e == hipSuccess) e = upload(ff, t->general, NGEN, &ff->dev.gen);
  if (e == hipSuccess) e = upload(ff, t->atom, nt * AP_COUNT, &ff->dev.atom);
  if (e == hipSuccess) e = upload(ff, t->pair, nt * nt * PP_COUNT, &ff->dev.pair);
  if (e == hipSuccess) e = upload(ff, t->angle_offsets, nt * nt * nt + 1, &ff->dev.ang_off);
  if (e == hipSuccess) e = upload(ff, t->angle_rows, nang * ANG_W, &ff->dev.ang);
  if (e == hipSuccess) e = upload(ff, t->torsion, nt * nt * nt * nt * TOR_W, &ff->dev.tor);
  if (e == hipSuccess) e = upload(ff, t->torsion_on, nt * nt * nt * nt, &ff->dev.tor_on);
  if (e == hipSuccess) {
    // Which type pairs (j, k) and triples (i, j, k), (j, k, l) have any torsion parameters: the torsion
    // kernel skips middle bonds and bond loops that no parameter row can match.
    std::vector<unsigned char> mid(nt * nt, 0), left(nt * nt * nt, 0), right(nt * nt * nt, 0);
    for (size_t i = 0; i < nt; ++i)
      for (size_t j = 0; j < nt; ++j)
        for (size_t k = 0; k < nt; ++k)
          for (size_t l = 0; l < nt; ++l)
            if (t->torsion_on[((i * nt + j) * nt + k) * nt + l]) {
              mid[j * nt + k] = 1;
              left[(i * nt + j) * nt + k] = 1;
              right[(j * nt + k) * nt + l] = 1;
            }
    e = upload(ff, mid.data(), nt * nt, &ff->dev.tor_mid);
    if (e == hipSuccess) e = upload(ff, left.data(), nt * nt * nt, &ff->dev.tor_left);
    if (e == hipSuccess) e = upload(ff, right.data(), nt * nt * nt, &ff->dev.tor_right);
  }
  if (e == hipSuccess) e = upload(ff, t->hbond, nt * nt * nt * HB_W, &ff->dev.hb);
  if (e == hipSuccess) e = upload(ff, t->hbond_on, nt * nt * nt, &ff->dev.hb_on);
  if (e == hipSuccess) {
    // per-pair constant of the shielded vdW distance (r^p + (1/gamma_w)^p)^(1/p)
    std::vector<double> b(nt * nt);
    for (size_t k = 0; k < nt * nt; ++k) b[k] = std::pow(1.0 / t->pair[k * PP_COUNT + PP_GAMMAW], t->general[28]);
    e = upload(ff, b.data(), nt * nt, &ff->dev.vdw_b);
  }
  if (e == hipSuccess) {
    // per-pair ratios of the van der Waals and inner-wall terms, so the edge
    // loop multiplies instead of dividing
    std::vector<double> c(4 * nt * nt);
    for (size_t k = 0; k < nt * nt; ++k) {
      const double* pp = t->pair + k * PP_COUNT;
      c[4 * k] = 1.0 / pp[PP_RVDW];
      c[4 * k + 1] = -0.5 * pp[PP_ALPHA] / pp[PP_RVDW];
      c[4 * k + 2] = 1.0 / pp[PP_RCORE];
      c[4 * k + 3] = -pp[PP_ACORE] / pp[PP_RCORE];
    }
    e = upload(ff, c.data(), 4 * nt * nt, &ff->dev.vdw_c);
  }
  if (e != hipSuccess) {
    xreac_ff_destroy(ff);
    return e == hipErrorOutOfMemory ? XREAC_NOMEM : XREAC_DEVICE;
  }
  *out = ff;
  return XREAC_OK;
}

void xreac_ff_destroy(xreac_ff* ff) {
  if (!ff) return;
  for (void* p : ff->allocations) (void)hipFree(p);
  delete ff;
}

int xreac_ctx_create(int device, void* stream, xreac_ctx** out) {
  if (!out) return XREAC_INVALID;
  if (device < 0 || device >= kMaxDevices) return XREAC_INVALID;
  if (hipSetDevice(device) != hipSuccess) return XREAC_DEVICE;
  auto* c = new xreac_ctx();
  c->device = device;
  c->stream = static_cast<hipStream_t>(stream);
  // Everything an evaluation needs besides device workspace is made here, so
  // that xreac_evaluate_static can run without allocating: the side stream
  // (default priority) and its events (ordering between two streams of one
  // device only: no system-scope fence), the pinned QEq poll slots and a
  // pinned control-block mirror, and this device's scratch address.
  bool ok = hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking) == hipSuccess &&
            hipEventCreateWithFlags(&c->bonds_ready, hipEventDisableTiming | hipEventDisableSystemFence) ==
                hipSuccess &&
            hipEventCreateWithFlags(&c->side_done, hipEventDisableTiming | hipEventDisableSystemFence) == hipSuccess;
  for (hipEvent_t& ev : c->poll_ev) ok = ok && hipEventCreateWithFlags(&ev, hipEventDisableTiming) == hipSuccess;
  ok = ok && hipHostMalloc(reinterpret_cast<void**>(&c->poll_host), 16 * sizeof(int)) == hipSuccess;
  const size_t ctl_bytes = 64 * 1024;
  if (ok && hipHostMalloc(reinterpret_cast<void**>(&c->es.ctl_host), ctl_bytes) == hipSuccess)
    c->es.ctl_host_bytes = ctl_bytes;
  else ok = false;
  if (ok && !g_scratch_ptr[device]) {
    void* p = nullptr;
    ok = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) == hipSuccess && p != nullptr;
    if (ok) g_scratch_ptr[device] = p;
  }
  c->scratch = g_scratch_ptr[device];
  if (!ok) {
    xreac_ctx_destroy(c);
    return XREAC_DEVICE;
  }
  *out = c;
  return XREAC_OK;
}

void xreac_ctx_destroy(xreac_ctx* c) {
  if (!c) return;
  if (c->side) {
    (void)hipStreamSynchronize(c->side);
    (void)hipStreamDestroy(c->side);
  }
  if (c->bonds_ready) (void)hipEventDestroy(c->bonds_ready);
  if (c->side_done) (void)hipEventDestroy(c->side_done);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (hipEvent_t ev : c->poll_ev)
    if (ev) (void)hipEventDestroy(ev);
  if (c->poll_host) (void)hipHostFree(c->poll_host);
  delete c;
}

const char* xreac_error_message(const xreac_ctx* c) { return c ? c->message.c_str() : "null context"; }

}  // extern "C"

namespace {

// Internal form of xreac_batch / xreac_batch_result.
struct BatchIn {
  int ns;
  const int32_t* offsets;  // host [ns + 1]
  const double* x;         // device [N*3]
  const int32_t* types;    // device [N]
  const double* cells;     // host [ns*9] or null
  const int32_t* pbc;      // host [ns*3] or null
  const double* charges;   // host [ns] or null
  const int32_t* nbr_row_ptr = nullptr;  // optional supplied CSR (device)
  const int32_t* nbr_j = nullptr;
  const int32_t* nbr_shifts = nullptr;
  int64_t nbr_count = 0;
};

struct BatchOut {
  double* comps;     // host [ns*NCOMP]
  int32_t* status;   // host [ns]
  int32_t* iters;    // host [ns]
  double* forces;    // device
  double* charges;
  double* tbo;
  double* nlp;
  int32_t* bcounts;
  double* qeq_state;
  int64_t num_edges = 0, num_bonds = 0;
};

template <class T>
size_t carve_at(size_t& off, size_t count) {
  off = (off + 255) & ~size_t(255);
  const size_t at = off;
  off += count * sizeof(T);
  return at;
}

// Control block: small per-call state, mirrored in pinned host memory so
// one copy brings every per-system result back.
struct CtlLayout {
  size_t flags, serr, comps, bounds, mu, cg, cg2, sys, gate, total;
};

CtlLayout ctl_layout(int ns) {
  CtlLayout c{};
  size_t off = 0;
  c.flags = carve_at<unsigned>(off, 16);  // [0] buffer-overflow bit, int [1] edges, [2] bonds, [3] longest row,
                                          // [5], [6] skin edges and longest row, [8] torsion long-atom count
  c.serr = carve_at<unsigned>(off, ns);
  c.comps = carve_at<double>(off, (size_t)ns * NCH);  // channel sums; components on the host
  c.bounds = carve_at<double>(off, (size_t)ns * 6);
  c.mu = carve_at<double>(off, ns);
  c.cg = carve_at<char>(off, xr_cg_state_bytes(ns));
  c.cg2 = carve_at<char>(off, xr_cg_state_bytes(ns));  // charge-response adjoint solve
  c.gate = carve_at<unsigned>(off, ns);  // skin reuse: 1 where the system's skin list was rebuilt
  c.sys = carve_at<SysDev>(off, ns);     // last: the copy back covers everything before it
  c.total = (off + 255) & ~size_t(255);
  return c;
}

__global__ void k_fill_systems(Systems S, int* __restrict__ asys) {
  const SysDev& sd = S.sys[blockIdx.x];
  for (int i = sd.atom0 + threadIdx.x; i < sd.atom0 + sd.n; i += blockDim.x) asys[i] = blockIdx.x;
}

// Bin grid of one system (rc/2-thick bins, searched two deep when periodic).
// The nonperiodic directions depend on the coordinate bounds and are fitted
// on the device (xr_bin_grid); here they get placeholders.
void bin_geometry(SysDev& sd, const double* height, double rcb) {
  BinGeom& g = sd.g;
  for (int k = 0; k < 3; ++k) {
    g.pbc[k] = sd.cell.pbc[k];
    if (sd.cell.pbc[k]) {
      int nbk = (int)std::floor(2.0 * height[k] / rcb);
      nbk = std::max(1, std::min(nbk, 128));
      g.nb[k] = nbk;
      g.w[k] = 1.0 / nbk;
      g.lo[k] = 0.0;
      const double h = height[k] / nbk;
      g.reach[k] = h >= rcb ? 1 : (int)std::floor(rcb / h) + 1;
    } else {
      g.nb[k] = 1;
      g.w[k] = 1.0;
      g.lo[k] = 0.0;
      g.reach[k] = 1;
    }
    g.thick[k] = g.w[k] * height[k];
  }
  // Gershgorin bound on the largest eigenvalue of the Gram matrix of the unit
  // plane normals (columns of the inverse cell): a displacement v has
  // sum_k (v . n_k)^2 <= kappa |v|^2, which lets the search prune bins by
  // their summed squared slab distances for any cell shape.
  double n[3][3];
  for (int k = 0; k < 3; ++k) {
    const double c0 = sd.cell.inv[k], c1 = sd.cell.inv[3 + k], c2 = sd.cell.inv[6 + k];
    const double len = std::sqrt(c0 * c0 + c1 * c1 + c2 * c2);
    n[k][0] = c0 / len;
    n[k][1] = c1 / len;
    n[k][2] = c2 / len;
  }
  double kappa = 1.0;
  for (int k = 0; k < 3; ++k) {
    double row = 0.0;
    for (int l = 0; l < 3; ++l)
      row += l == k ? 1.0 : std::fabs(n[k][0] * n[l][0] + n[k][1] * n[l][1] + n[k][2] * n[l][2]);
    kappa = std::max(kappa, row);
  }
  g.kappa = kappa * (1.0 + 1e-12);
}

// Workspace of `need` bytes: the context's grow-only buffer, or the next piece
// of the arena when the evaluation runs in static scratch.
hipError_t workspace(Arena* arena, Buffer& buf, size_t need, void** out) {
  if (!arena) {
    hipError_t e = buf.ensure(need);
    *out = buf.ptr;
    return e;
  }
  const size_t off = (arena->used + 255) & ~size_t(255);
  arena->wanted = off + need;
  if (off + need > arena->cap) return hipErrorOutOfMemory;
  *out = arena->base + off;
  arena->used = off + need;
  return hipSuccess;
}

// What the host side of an evaluation needs once the device is done.
struct Finish {
  CtlLayout L{};
  int ns = 0, n = 0;
  bool given_q = false, full = false, supplied = false;
  bool fixed = false;  // a fixed number of CG launches (a recorded step)
  double skin = 0.0;
  int64_t m_cap = 0, nbond = 0, nbr_count = 0;
  int* row_ptr = nullptr;
  int* nbr_j = nullptr;
  int* nbr_s = nullptr;
  int* bond_ptr = nullptr;
  int* bond_i = nullptr;
  int* bond_j = nullptr;
  int* bond_code = nullptr;
  double* bond_BO = nullptr;
};

// Recording an evaluation into a graph (stream capture): no read-back, a
// fixed number of CG launches per solve, and no buffer may be (re)allocated
// while the stream captures; a workspace that would grow is named instead
// and the recording is abandoned, to be grown and recorded again.
struct Capture {
  int cg_launches = 0;
  Finish fin;
  Buffer* grow = nullptr;
  size_t grow_bytes = 0;
  size_t grow_pinned = 0;  // bytes of pinned control block needed
};

bool cg_final(int state) { return state == CG_DONE || state == CG_STALLED || state == CG_FAILED; }

// Host side of an evaluation once the device is done with it: overflow
// detection, per-system status, components and QEq iterations, and the
// sizing history the next evaluation's buffers are taken from.
int finish_eval(xreac_ctx* ctx, EvalState& st, const Finish& f, BatchOut& out, bool* overflow) {
  *overflow = false;
  char* hc = st.ctl_host;
  const CtlLayout& L = f.L;
  const int ns = f.ns;
  const unsigned* h_flags = reinterpret_cast<const unsigned*>(hc + L.flags);
  const int* host_counts = reinterpret_cast<const int*>(h_flags) + 1;
  const int64_t m = host_counts[0], nb_total = host_counts[1];
  if (f.supplied && m != f.nbr_count)
    return fail(ctx, XREAC_INVALID, "neighbor_row_ptr does not match neighbor_count");
  if (m > f.m_cap || nb_total > f.nbond || (h_flags[0] & XR_ERR_CAPACITY)) {
    *overflow = true;
    return XREAC_OK;
  }
  const unsigned* h_serr = reinterpret_cast<const unsigned*>(hc + L.serr);
  const double* h_comps = reinterpret_cast<const double*>(hc + L.comps);
  const CgSys* h_cg = xr_cg_systems(hc + L.cg);
  const CgSys* h_cg2 = xr_cg_systems(hc + L.cg2);
  // With a fixed launch count a system still running when the launches ran
  // out is not a failure: the step is repeated with read-backs.
  if (f.fixed)
    for (int q = 0; q < ns; ++q)
      if (!h_serr[q] && ((!f.given_q && !cg_final(h_cg[q].ctl[0].done)) || (f.full && !cg_final(h_cg2[q].ctl[0].done)))) {
        *overflow = true;
        return XREAC_OK;
      }
  ctx->sys_messages.assign(ns, std::string());
  int launches = 0;
  for (int q = 0; q < ns; ++q) {
    out.iters[q] = f.given_q ? 0 : h_cg[q].iters;
    if (!f.given_q && h_cg[q].ctl[0].done == CG_DONE) launches = std::max(launches, h_cg[q].ks + 1);
    if (f.full && h_cg2[q].ctl[0].done == CG_DONE) launches = std::max(launches, h_cg2[q].ks + 1);
    xr_components(h_comps + (size_t)q * NCH, out.comps + (size_t)q * NCOMP);
    if (h_serr[q]) {
      out.status[q] = XREAC_INVALID;
      ctx->sys_messages[q] = error_text(h_serr[q]);
    } else if ((!f.given_q && h_cg[q].ctl[0].done != CG_DONE) || (f.full && h_cg2[q].ctl[0].done != CG_DONE)) {
      out.status[q] = XREAC_INVALID;
      // calculator.py:149-150 raises this text when the dense solve fails
      ctx->sys_messages[q] =
          "Singular QEq system; check the geometry and parameters (the QEq conjugate gradient did not converge)";
    } else {
      out.status[q] = XREAC_OK;
    }
  }
  if (f.skin > 0.0) {
    const unsigned* h_gate = reinterpret_cast<const unsigned*>(hc + L.gate);
    bool clean = true;
    for (int q = 0; q < ns; ++q) {
      ctx->neighbor_list_builds += h_gate[q] ? 1 : 0;
      clean = clean && h_serr[q] == 0;
    }
    st.skin_valid = clean;  // a failed system may have left its kept rows incomplete
  }
  out.num_edges = m;
  out.num_bonds = nb_total;
  st.m_per_atom = (double)m / f.n;
  st.b_per_atom = (double)nb_total / f.n;
  if (!f.supplied) st.max_row = host_counts[2];
  st.last_cg_launches = launches;
  if (launches > 0) {
    st.cg_history[st.cg_history_next] = launches;
    st.cg_history_next = (st.cg_history_next + 1) % 8;
  }

  st.n = f.n;
  st.m = m;
  st.nb = nb_total;
  st.row_ptr = f.row_ptr;
  st.nbr_j = f.nbr_j;
  st.nbr_s = f.nbr_s;
  st.bond_ptr = f.bond_ptr;
  st.bond_i = f.bond_i;
  st.bond_j = f.bond_j;
  st.bond_code = f.bond_code;
  st.bond_BO = f.bond_BO;
  return XREAC_OK;
}

// XREAC_BOND_CAPACITY_SCALE (default 1, read once per process) scales the
// bond capacity an evaluation sizes from the previous call's bonds per atom,
// slack included; below 1 such first attempts overflow their bond buffers on
// the second stream and are repeated with exact sizes, which tests exercise.
double bond_capacity_scale() {
  static const double scale = [] {
    const char* v = std::getenv("XREAC_BOND_CAPACITY_SCALE");
    const double d = v ? std::atof(v) : 1.0;
    return d > 0.0 && d < 1.0 ? d : 1.0;
  }();
  return scale;
}

// One attempt.  With exact == false, edge and bond buffers are sized from the
// previous evaluation's per-atom counts and no mid-pipeline read-back happens;
// if a count exceeds its buffer, *overflow is set and the caller repeats the
// evaluation with exact == true (sizes read back before allocation).
int evaluate_once(xreac_ctx* ctx, EvalState& st, const xreac_ff* ffh, const BatchIn& in, const xreac_options* opt,
                  BatchOut& out, bool exact, bool* overflow, Arena* arena, Capture* cap = nullptr) {
  *overflow = false;
  if (cap && (exact || arena)) return fail(ctx, XREAC_INVALID, "internal error: a recorded evaluation is sized ahead");
  if (arena) {
    arena->used = 0;
    arena->wanted = 0;
    exact = true;  // sizes read back before carving: no hints, no second attempt
  }
  const int ns = in.ns;
  const int n = in.offsets[ns];
  HIPCHK(hipSetDevice(ctx->device), "hipSetDevice");
  hipStream_t s = ctx->stream;
  const FFDev& ff = ffh->dev;
  // Timers are resolved after every synchronisation that may end the call.
  struct ActiveProfiler {
    XrProfiler* p;
    explicit ActiveProfiler(XrProfiler* q) : p(q) { xr_active_profiler = q; }
    ~ActiveProfiler() {
      // An early return leaves recorded events unresolved; drop them.
      if (p) {
        p->pending.clear();
        p->used = 0;
      }
      xr_active_profiler = nullptr;
    }
  } active(arena || cap ? nullptr : &ctx->prof);  // no timer events in static scratch or in a recording
  const double rc = ffh->cutoff;
  const double bond_cut = std::min(5.0, rc);
  // ase.py:90-93 (the same check and text)
  if (!(opt->neighbor_skin >= 0.0) || !std::isfinite(opt->neighbor_skin))
    return fail(ctx, XREAC_INVALID, "neighbor_skin must be a finite nonnegative number");
  // Skin reuse applies to lists built here (not to supplied ones).  As ASE's neighbor_skin (ase.py:39-40,
  // 112-124): a per-atom displacement allowance, so the kept rows reach rc + 2 skin, and bins are that wide.
  const double skin =
      (in.nbr_row_ptr == nullptr && arena == nullptr && cap == nullptr && opt->neighbor_skin > 0.0)
          ? opt->neighbor_skin
          : 0.0;
  const double rc_keep = rc + 2.0 * skin;
  const double rcb = rc_keep * (1 + 1e-8) + 1e-12;
  const bool mixed = opt->precision == XREAC_PRECISION_MIXED;
  const bool given_q = opt->fixed_charges != 0;
  const bool full = opt->full_derivative != 0 && !given_q;
  const int lammps = full ? 0 : opt->lammps_angles;

  // ------------------------------------------------------- control block
  const CtlLayout L = ctl_layout(ns);
  // The arena's workspaces are checked against its capacity as they are taken.
#define XR_WORKSPACE(buf, need, out, what)                                                                    \
  do {                                                                                                        \
    if (cap && (size_t)(need) > (buf).bytes) {                                                                \
      cap->grow = &(buf);                                                                                     \
      cap->grow_bytes = (need);                                                                               \
      return XREAC_OK;                                                                                        \
    }                                                                                                         \
    hipError_t e_ = workspace(arena, (buf), (need), (out));                                                   \
    if (e_ != hipSuccess && arena)                                                                            \
      return fail(ctx, XREAC_NOMEM,                                                                           \
                  "xreac_evaluate_static: the system needs more than the " +                                 \
                      std::to_string(XREAC_STATIC_SCRATCH_BYTES >> 20) +                                      \
                      " MiB of static scratch (at least " + std::to_string(arena->wanted >> 20) +            \
                      " MiB); evaluate it with xreac_evaluate");                                              \
    if (e_ != hipSuccess) return hip_fail(ctx, e_, what);                                                     \
  } while (0)
  void* ctl_dev = nullptr;
  XR_WORKSPACE(st.ctl_buf, L.total, &ctl_dev, "control block");
  if (arena && st.ctl_host_bytes < L.total)
    return fail(ctx, XREAC_NOMEM, "xreac_evaluate_static: the pinned control block made at context creation is "
                                  "too small for this call");
  if (cap && st.ctl_host_bytes < L.total) {
    cap->grow_pinned = L.total;
    return XREAC_OK;
  }
  if (st.ctl_host_bytes < L.total) {
    if (st.ctl_host) (void)hipHostFree(st.ctl_host);
    st.ctl_host = nullptr;
    st.ctl_host_bytes = 0;
    const size_t want = L.total + L.total / 2;
    HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&st.ctl_host), want), "pinned control block");
    st.ctl_host_bytes = want;
  }
  char* dc = static_cast<char*>(ctl_dev);
  char* hc = st.ctl_host;
  unsigned* d_flags = reinterpret_cast<unsigned*>(dc + L.flags);
  SysDev* hsys = reinterpret_cast<SysDev*>(hc + L.sys);
  std::vector<double> heights(3 * (size_t)ns);
  int nd = 0, nu = 0, nd_small = 0, max_sys = 0;
  bool need_bounds = false;
  for (int q = 0; q < ns; ++q) {
    SysDev sd;
    std::memset(&sd, 0, sizeof(sd));
    sd.atom0 = in.offsets[q];
    sd.n = in.offsets[q + 1] - in.offsets[q];
    const double* cp = in.cells ? in.cells + 9 * (size_t)q : nullptr;
    int pb[3];
    for (int k = 0; k < 3; ++k) pb[k] = cp ? (in.pbc ? (in.pbc[3 * q + k] != 0) : 1) : 0;
    const bool periodic = cp && (pb[0] || pb[1] || pb[2]);
    CellDev& cell = sd.cell;
    if (periodic) {
      std::memcpy(cell.L, cp, sizeof(cell.L));
    } else {
      for (int k = 0; k < 9; ++k) cell.L[k] = (k % 4 == 0) ? 1.0 : 0.0;
    }
    invert3(cell.L, cell.inv);
    for (int k = 0; k < 3; ++k) cell.pbc[k] = periodic ? pb[k] : 0;
    cell.periodic = periodic ? 1 : 0;
    for (int k = 0; k < 3; ++k) {
      const double c0 = cell.inv[k], c1 = cell.inv[3 + k], c2 = cell.inv[6 + k];
      heights[3 * q + k] = 1.0 / std::sqrt(c0 * c0 + c1 * c1 + c2 * c2);
      need_bounds |= !cell.pbc[k];
    }
    max_sys = std::max(max_sys, sd.n);
    xr_cg_blocks(sd.n, &sd.ndir, &sd.nupd, &sd.chunk);
    sd.upd0 = nu;
    nd += sd.ndir;
    nu += sd.nupd;
    if (xr_cg_small(sd.n)) nd_small += sd.ndir;
    sd.total_charge = in.charges ? in.charges[q] : 0.0;
    hsys[q] = sd;
  }
  // CG iteration blocks: small systems' first, then large ones', each in batch order
  for (int q = 0, ds = 0, dl = nd_small; q < ns; ++q) {
    int& d = xr_cg_small(hsys[q].n) ? ds : dl;
    hsys[q].dir0 = d;
    d += hsys[q].ndir;
  }

  // ------------------------------------------------------ atom workspace
  size_t off = 0;
  const size_t scan_ints = xr_scan_scratch_ints(std::max(n, 1 << 16));
  carve_size<int>(off, n);               // system of each atom
  carve_size<int>(off, 3 * (size_t)n);   // home
  carve_size<int>(off, n);               // bin
  carve_size<int>(off, n);               // counts
  carve_size<int>(off, n + 1);           // row_ptr
  carve_size<int>(off, n);               // bond_count
  carve_size<int>(off, n + 1);           // bond_ptr
  carve_size<int>(off, scan_ints);       // scan scratch
  carve_size<char>(off, xr_bin_record_bytes() * n);  // bin-ordered atom records
  carve_size<double2>(off, 10 * (size_t)n);  // CG vectors: x, r, u, w, z, q, s, p and two m (pairs)
  carve_size<float2>(off, 2 * (size_t)n);    // FP32 copies of the two m (mixed precision)
  carve_size<int2>(off, nu);                 // CG chunk ranges
  carve_size<CgBlock>(off, nd);              // CG iteration block descriptors and row pointers
  carve_size<double>(off, xr_cg_block_doubles(n));  // block-Jacobi inverse rows
  carve_size<int2>(off, n);              // block placement
  carve_size<unsigned char>(off, n);     // block join flags
  carve_size<double>(off, n);            // eta
  carve_size<double>(off, xr_cg_partials(nd, nu));
  carve_size<double>(off, (size_t)kAtomScalars * n);  // atom scalars
  carve_size<int>(off, n);               // bond counts (>0.3)
  carve_size<int>(off, n);               // H-bond candidates per row
  carve_size<int>(off, n);               // atoms with more bonds than a torsion lane group takes
  carve_size<double>(off, 6 * (size_t)n);   // F_nb, F_hb
  carve_size<double>(off, (size_t)NCH * n);  // per-atom energy channels
  carve_size<double2>(off, full ? n : 0);    // charge-response right-hand sides (dE/dq, -1)
  carve_size<double>(off, full ? n : 0);     // charge-response adjoint lambda
  void* atom_ws = nullptr;
  XR_WORKSPACE(st.atom_buf, off + 4096, &atom_ws, "atom workspace");
  Carver cv{static_cast<char*>(atom_ws)};
  int* asys = cv.take<int>(n);
  int* home = cv.take<int>(3 * (size_t)n);
  int* bin = cv.take<int>(n);
  int* counts = cv.take<int>(n);
  int* row_ptr = cv.take<int>(n + 1);
  int* bond_count = cv.take<int>(n);
  int* bond_ptr = cv.take<int>(n + 1);
  int* scan_tmp = cv.take<int>(scan_ints);
  void* binned = cv.take<char>(xr_bin_record_bytes() * n);
  double2* cgv = cv.take<double2>(10 * (size_t)n);
  float2* cgq = cv.take<float2>(2 * (size_t)n);
  int2* cg_chunks = cv.take<int2>(nu);
  CgBlock* cg_blocks = cv.take<CgBlock>(nd);
  double* minv = cv.take<double>(xr_cg_block_doubles(n));
  int2* binfo = cv.take<int2>(n);
  unsigned char* join = cv.take<unsigned char>(n);
  double* eta = cv.take<double>(n);
  double* cg_part = cv.take<double>(xr_cg_partials(nd, nu));
  double* scal = cv.take<double>((size_t)kAtomScalars * n);
  double* const scal_end = scal + (size_t)kAtomScalars * n;
  int* bcount03 = cv.take<int>(n);
  int* hb_count = cv.take<int>(n);
  int* tor_long = cv.take<int>(n);
  double* fbuf = cv.take<double>(6 * (size_t)n);
  double* energies = cv.take<double>((size_t)NCH * n);
  double2* dEdq = full ? cv.take<double2>(n) : nullptr;
  double* lam = full ? cv.take<double>(n) : nullptr;

  Systems S{reinterpret_cast<const SysDev*>(dc + L.sys), asys, reinterpret_cast<unsigned*>(dc + L.serr), ns};

  // ------------------------------------------------------- bin grids
  // Periodic directions are fixed by the cell.  A system with a nonperiodic
  // direction gets a budget of bins (4 per atom + 64, whole multiples of its
  // periodic bin count), and its grid is fitted to its coordinate bounds on
  // the device, so the host never waits for the bounds.
  long long nbins = 0;
  for (int q = 0; q < ns; ++q) {
    SysDev& sd = hsys[q];
    bin_geometry(sd, &heights[3 * (size_t)q], rcb);
    const long long periodic = (long long)sd.g.nb[0] * sd.g.nb[1] * sd.g.nb[2];  // placeholders count 1
    long long budget = periodic;
    if (!(sd.cell.pbc[0] && sd.cell.pbc[1] && sd.cell.pbc[2])) {
      const long long np = std::max(1LL, std::min(128LL * 128 * 128, (4LL * sd.n + 64) / periodic));
      budget = periodic * np;
    }
    sd.bin0 = (int)nbins;
    sd.bin_budget = (int)std::min<long long>(budget, INT32_MAX / 2);
    nbins += budget;
    if (nbins > INT32_MAX / 2) return fail(ctx, XREAC_INVALID, "too many cell-list bins in one batch");
  }
  HIPCHK(hipMemcpyAsync(dc + L.sys, hc + L.sys, sizeof(SysDev) * ns, hipMemcpyHostToDevice, s), "system table");

  off = 0;
  carve_size<int>(off, nbins + 1);
  carve_size<int>(off, nbins + 1);
  carve_size<int>(off, nbins + 1);
  carve_size<int>(off, xr_scan_scratch_ints((int)nbins));
  void* bin_ws = nullptr;
  XR_WORKSPACE(st.bin_buf, off + 4096, &bin_ws, "bin workspace");
  Carver cb{static_cast<char*>(bin_ws)};
  int* bin_count = cb.take<int>(nbins + 1);
  int* bin_start = cb.take<int>(nbins + 1);
  int* bin_fill = cb.take<int>(nbins + 1);
  int* bin_scan = cb.take<int>(xr_scan_scratch_ints((int)nbins));

  Atoms at{};
  at.n = n;
  at.S = S;
  at.x = in.x;
  at.type = in.types;
  double* sp = scal;
  auto nxt = [&]() {
    double* p = sp;
    sp += n;
    return p;
  };
  at.totp = nxt();
  at.total = nxt();
  at.d = nxt();
  at.db = nxt();
  at.nlp = nxt();
  at.vlpex = nxt();
  at.dlp = nxt();
  at.dlpt = nxt();
  at.sbo = nxt();
  at.sbo2 = nxt();
  at.prod = nxt();
  at.ov2 = nxt();
  at.A_total = nxt();
  at.A_d = nxt();
  at.A_db = nxt();
  at.A_nlp = nxt();
  at.A_dlpt = nxt();
  at.A_dlp = nxt();
  at.A_sbo2 = nxt();
  at.a_ov2 = nxt();
  at.a_totp = nxt();
  at.bomax = nxt();
  if (sp > scal_end) return fail(ctx, XREAC_INVALID, "internal error: per-atom workspace layout");
  // Per-atom properties the caller asked for are computed in its buffers directly.
  if (out.tbo) at.total = out.tbo;
  if (out.nlp) at.nlp = out.nlp;
  at.bond_count = out.bcounts ? out.bcounts : bcount03;
  at.F_nb = fbuf;
  at.F_hb = fbuf + 3 * (size_t)n;
  at.q = out.charges;
  EnergyAcc acc{energies, n};

  int* dev_counts = reinterpret_cast<int*>(d_flags) + 1;  // [0] total edges, [1] total bonds, [2] longest row
  {
    ZeroRanges z{};
    auto add = [&](void* p, size_t bytes) {
      z.ptr[z.count] = static_cast<unsigned*>(p);
      z.words[z.count] = (long long)(bytes / 4);
      ++z.count;
    };
    add(d_flags, 16 * sizeof(unsigned));
    add(dc + L.serr, sizeof(unsigned) * ns);
    add(dc + L.cg, xr_cg_state_bytes(ns) & ~size_t(3));
    add(dc + L.cg2, xr_cg_state_bytes(ns) & ~size_t(3));
    add(bin_count, sizeof(int) * (nbins + 1));
    add(bin_fill, sizeof(int) * (nbins + 1));
    add(energies, sizeof(double) * NCH * (size_t)n);
    add(at.F_hb, sizeof(double) * 3 * (size_t)n);
    hipLaunchKernelGGL(k_zero_ranges, dim3(256), dim3(256), 0, s, z);
    HIPCHK(hipGetLastError(), "zero");
  }
  hipLaunchKernelGGL(k_fill_systems, dim3(ns), dim3(256), 0, s, S, asys);
  HIPCHK(hipGetLastError(), "system index");
  if (need_bounds) {
    HIPCHK(xr_coord_bounds(S, in.x, reinterpret_cast<double*>(dc + L.bounds), s), "bounds");
    HIPCHK(xr_bin_grid(reinterpret_cast<SysDev*>(dc + L.sys), ns, reinterpret_cast<const double*>(dc + L.bounds),
                       rcb, reinterpret_cast<unsigned*>(dc + L.serr), s),
           "bin grid");
  }
  HIPCHK(xr_bin_atoms(n, in.x, S, home, bin, bin_count, s), "bin");
  HIPCHK(xr_exclusive_scan(bin_count, (int)nbins, bin_start, bin_scan, s), "bin scan");
  HIPCHK(xr_bin_scatter(n, in.x, S, home, bin, bin_start, bin_fill, binned, s), "bin scatter");

  // --------------------------------------------------------- neighbor rows
  // The fill writes each row, sorted, to a slab of fixed stride and its length
  // to counts; one scan gives the row pointers and the edge pass moves the rows
  // into place.  Buffers are sized from the last evaluation; an exact
  // evaluation (the first, or after an overflow) counts the rows first.  A
  // supplied list gets the count pass too: it checks every pair for
  // coincident atoms, as validate_input does whatever list is supplied.
  const bool supplied = in.nbr_row_ptr != nullptr;
  int max_row = st.max_row;
  int skin_max_row = 0;
  int64_t m_cap = 0;
  if (supplied) {
    HIPCHK(xr_neighbor_count(n, in.x, S, home, bin, bin_start, binned, rc, -1, counts, s), "neighbor count");
    m_cap = in.nbr_count;
    HIPCHK(hipMemcpyAsync(row_ptr, in.nbr_row_ptr, sizeof(int) * (n + 1), hipMemcpyDeviceToDevice, s), "copy");
    // clamped to [0, supplied length], so a malformed pointer array cannot reach outside the list;
    // a first pointer other than 0 or a decreasing pointer fails the system
    HIPCHK(xr_clamp_supplied(row_ptr, n, (int)m_cap, dev_counts, d_flags, S, s), "clamp");
  } else {
    const double slack = 1.10;
    m_cap = exact || st.m_per_atom <= 0 ? 0 : (int64_t)std::ceil(st.m_per_atom * n * slack) + 1024;
    exact = exact || m_cap == 0;
    if (cap && exact) return fail(ctx, XREAC_INVALID, "recording a step needs an earlier evaluation to size it");
    if (m_cap > INT32_MAX - 1) m_cap = INT32_MAX - 1;
    if (exact) {
      if (skin > 0.0) {  // the longest row within rc + 2 skin sizes the kept rows
        HIPCHK(xr_neighbor_count(n, in.x, S, home, bin, bin_start, binned, rc_keep, 2048, counts, s),
               "neighbor count");
        HIPCHK(xr_scan_clamped(counts, n, row_ptr, scan_tmp, INT32_MAX, dev_counts + 4, d_flags, dev_counts + 5, s),
               "neighbor scan");
      }
      HIPCHK(xr_neighbor_count(n, in.x, S, home, bin, bin_start, binned, rc, 2048, counts, s), "neighbor count");
      HIPCHK(xr_scan_clamped(counts, n, row_ptr, scan_tmp, INT32_MAX, dev_counts, d_flags, dev_counts + 2, s),
             "neighbor scan");
      int h[6] = {0, 0, 0, 0, 0, 0};
      HIPCHK(hipMemcpyAsync(h, dev_counts, sizeof(h), hipMemcpyDeviceToHost, s), "copy");
      HIPCHK(hipStreamSynchronize(s), "neighbor sync");
      m_cap = h[0];
      max_row = h[2];
      skin_max_row = h[5];
    }
  }

  off = 0;
  // The arrays the CG sweep reads (columns and H values) hold one spare entry
  // past the list: its batch loads read entry row_ptr[i] of an empty row i.
  carve_size<int>(off, m_cap + 1);
  carve_size<int>(off, m_cap);
  carve_size<double>(off, m_cap + 1);
  carve_size<float>(off, mixed ? m_cap + 1 : 0);
  // 2-byte CG columns (relative to the system's first atom): every system has at most 65,536 atoms
  const bool narrow = max_sys <= 65536;
  carve_size<unsigned short>(off, narrow ? m_cap + 1 : 0);
  carve_size<unsigned char>(off, m_cap);
  carve_size<int2>(off, m_cap);  // H-bond candidates, at the start of each row's range
  void* edge_ws = nullptr;
  XR_WORKSPACE(st.edge_buf, off + 4096, &edge_ws, "edge workspace");
  Carver ce{static_cast<char*>(edge_ws)};
  int* nbr_j = ce.take<int>(m_cap + 1);
  int* nbr_s = ce.take<int>(m_cap);
  double* hval = ce.take<double>(m_cap + 1);
  float* hval32 = mixed ? ce.take<float>(m_cap + 1) : nullptr;
  unsigned short* col16 = narrow ? ce.take<unsigned short>(m_cap + 1) : nullptr;
  unsigned char* bflag = ce.take<unsigned char>(m_cap);
  int2* hb_cand = ce.take<int2>(m_cap);
  void* slab = nullptr;
  int slab_stride = 0;
  if (supplied) {
    HIPCHK(xr_pack_supplied(n, S, row_ptr, in.nbr_j, in.nbr_shifts, nbr_j, nbr_s, s), "supplied neighbors");
  } else {
    // The fill instance (and slab stride) is sized from the longest row: known
    // in an exact evaluation, else the last evaluation's plus a margin.  A
    // longer row makes the instance raise the capacity bit and the evaluation
    // is repeated exactly.
    const int hint = exact ? max_row : (max_row > 0 ? max_row + 64 : 2048);
    slab_stride = xr_fill_stride(hint);
    XR_WORKSPACE(st.slab_buf, sizeof(unsigned long long) * (size_t)slab_stride * n + 4096, &slab,
                 "neighbor slab");
    if (skin > 0.0) {
      // Skin reuse.  The kept rows' stride comes from an exact count, else stays as kept, else from the
      // longest rc row scaled by the volume ratio; a row longer than it repeats the call exactly.
      int hint_s;
      if (exact) {
        hint_s = skin_max_row;
      } else if (st.skin_stride > 0) {
        hint_s = st.skin_stride;
      } else {
        const double grow = std::pow(rc_keep / rc, 3.0);
        hint_s = max_row > 0 ? (int)std::min(2048.0, std::ceil(max_row * grow * 1.1) + 64) : 2048;
      }
      const int stride_s = xr_fill_stride(hint_s);
      SkinKey key;
      key.rc = rc;
      key.skin = skin;
      key.offsets.assign(in.offsets, in.offsets + ns + 1);
      if (in.cells) key.cells.assign(in.cells, in.cells + 9 * (size_t)ns);
      if (in.pbc) key.pbc.assign(in.pbc, in.pbc + 3 * (size_t)ns);
      bool force = exact || !st.skin_valid || stride_s != st.skin_stride || !(key == st.skin_key);
      const void* kept[4] = {st.skin_slab_buf.ptr, st.skin_counts_buf.ptr, st.skin_xref_buf.ptr,
                             st.skin_tref_buf.ptr};
      HIPCHK(st.skin_slab_buf.ensure(sizeof(unsigned long long) * (size_t)stride_s * n + 4096), "skin rows");
      HIPCHK(st.skin_counts_buf.ensure(sizeof(int) * (size_t)n + 256), "skin rows");
      HIPCHK(st.skin_xref_buf.ensure(sizeof(double) * 3 * (size_t)n + 256), "skin rows");
      HIPCHK(st.skin_tref_buf.ensure(sizeof(int) * (size_t)n + 256), "skin rows");
      force = force || kept[0] != st.skin_slab_buf.ptr || kept[1] != st.skin_counts_buf.ptr ||
              kept[2] != st.skin_xref_buf.ptr || kept[3] != st.skin_tref_buf.ptr;
      st.skin_valid = false;  // until this call completes cleanly
      st.skin_stride = stride_s;
      st.skin_key = std::move(key);
      unsigned* gate = reinterpret_cast<unsigned*>(dc + L.gate);
      int* skin_counts = static_cast<int*>(st.skin_counts_buf.ptr);
      double* xref = static_cast<double*>(st.skin_xref_buf.ptr);
      int* tref = static_cast<int*>(st.skin_tref_buf.ptr);
      // atoms that moved at most skin (less a rounding margin) keep every pair within rc in the kept rows
      HIPCHK(xr_skin_check(S, in.x, in.types, xref, tref, skin * (1.0 - 1e-9), force, gate, s), "skin check");
      HIPCHK(xr_neighbor_fill(n, in.x, S, home, bin, bin_start, binned, rc_keep, skin_counts,
                              st.skin_slab_buf.ptr, stride_s, d_flags, s, gate),
             "neighbor fill");
      HIPCHK(xr_skin_commit(n, S, in.x, in.types, xref, tref, gate, s), "skin commit");
      HIPCHK(xr_skin_filter(n, in.x, S, st.skin_slab_buf.ptr, stride_s, skin_counts, rc, counts, slab,
                            slab_stride, d_flags, s),
             "skin filter");
    } else {
      HIPCHK(xr_neighbor_fill(n, in.x, S, home, bin, bin_start, binned, rc, counts, slab, hint, d_flags, s),
             "neighbor fill");
    }
    // row pointers clamped to the edge buffer, total edges and longest row in one pass
    HIPCHK(xr_scan_clamped(counts, n, row_ptr, scan_tmp, exact ? INT32_MAX : (int)m_cap, dev_counts, d_flags,
                           dev_counts + 2, s),
           "neighbor scan");
  }

  // ------------------------------------------------------- edges and bonds
  HIPCHK(xr_edge_prep(n, in.x, in.types, S, ff, row_ptr, slab, slab_stride, nbr_j, nbr_s, col16, bond_cut, hval,
                      hval32, bflag, bond_count, join, hb_cand, hb_count, s),
         "edge pass");
  // The bond list (scan of the active-bond counts, compaction) feeds only the
  // bonded terms, so with buffers sized from the last evaluation it is built on
  // the second stream with them; an exact evaluation reads its size back first.
  const bool bonds_on_side = !exact;
  int64_t b_cap = 0;
  if (!exact) {
    const double scale = bond_capacity_scale();
    b_cap = (int64_t)std::ceil(st.b_per_atom * n * 1.25 * scale) + (int64_t)(1024 * scale);
    if (b_cap > m_cap) b_cap = m_cap;
  } else {
    HIPCHK(xr_scan_clamped(bond_count, n, bond_ptr, scan_tmp, INT32_MAX, dev_counts + 1, d_flags, nullptr, s),
           "bond scan");
    int nb_host = 0;
    HIPCHK(hipMemcpyAsync(&nb_host, dev_counts + 1, sizeof(int), hipMemcpyDeviceToHost, s), "copy");
    HIPCHK(hipStreamSynchronize(s), "bond sync");
    b_cap = nb_host;
  }
  const int64_t nbond = b_cap;

  off = 0;
  for (int k = 0; k < 5; ++k) carve_size<int>(off, nbond);  // edge, bi, bj, code, rev
  carve_size<double>(off, 3 * (size_t)nbond);               // vec
  carve_size<double>(off, 3 * (size_t)nbond);               // G
  carve_size<double>(off, (size_t)kBondScalars * nbond);    // per-bond scalars
  void* bond_ws = nullptr;
  XR_WORKSPACE(st.bond_buf, off + 4096, &bond_ws, "bond workspace");
#undef XR_WORKSPACE
  Carver cbd{static_cast<char*>(bond_ws)};
  Bonds bd{};
  bd.cap = (int)nbond;
  bd.nbp = dev_counts + 1;
  bd.ptr = bond_ptr;
  int* b_edge = cbd.take<int>(nbond);
  int* b_i = cbd.take<int>(nbond);
  bd.edge = b_edge;
  bd.bi = b_i;
  bd.bj = cbd.take<int>(nbond);
  bd.code = cbd.take<int>(nbond);
  bd.rev = cbd.take<int>(nbond);
  bd.vec = cbd.take<double>(3 * (size_t)nbond);
  bd.G = cbd.take<double>(3 * (size_t)nbond);
  double* bs = cbd.take<double>((size_t)kBondScalars * nbond);
  double* const bs_end = bs + (size_t)kBondScalars * nbond;
  auto bn = [&]() {
    double* p = bs;
    bs += nbond;
    return p;
  };
  bd.r = bn();
  bd.bop = bn();
  bd.pim = bn();
  bd.ppm = bn();
  bd.dsig = bn();
  bd.dpi = bn();
  bd.dpp = bn();
  bd.f1 = bn();
  bd.f4 = bn();
  bd.f5 = bn();
  bd.boc = bn();
  bd.pic = bn();
  bd.ppc = bn();
  bd.BO = bn();
  bd.PI = bn();
  bd.PP = bn();
  bd.SIG = bn();
  bd.A_BO = bn();
  bd.A_PI = bn();
  bd.A_PP = bn();
  bd.A_SIG = bn();
  bd.A_totfar = bn();
  bd.a_bop = bn();
  bd.a_pim = bn();
  bd.a_ppm = bn();
  bd.c_i = bn();
  bd.c_j = bn();
  if (bs > bs_end) return fail(ctx, XREAC_INVALID, "internal error: per-bond workspace layout");
  if (!bonds_on_side && nbond > 0) HIPCHK(xr_bond_fill(n, row_ptr, bflag, bond_ptr, (int)nbond, b_edge, b_i, s), "bond fill");

  // ------------------------------------------- bonded terms on the side stream
  // Bond orders, many-body terms and their adjoints depend on positions only;
  // they run on a second stream while QEq runs here, and the force assembly
  // waits for them.  Every return below joins the side stream first.
  if (!ctx->side) return fail(ctx, XREAC_DEVICE, "internal error: context without its side stream");
  hipStream_t s2 = ctx->side;
  HIPCHK(hipEventRecord(ctx->bonds_ready, s), "event");
  // Issued from inside the QEq solve once its first iterations are queued, so
  // the main stream does not sit idle while the host makes these launches.
  const char* side_what = "side stream";
  auto launch_side = [&]() -> hipError_t {
    hipError_t e = hipStreamWaitEvent(s2, ctx->bonds_ready, 0);
    if (e != hipSuccess) return e;
    if (bonds_on_side) {
      side_what = "bond scan";
      e = xr_scan_clamped(bond_count, n, bond_ptr, scan_tmp, (int)b_cap, dev_counts + 1, d_flags, nullptr, s2);
      if (e != hipSuccess) return e;
      side_what = "bond fill";
      if (nbond > 0 && (e = xr_bond_fill(n, row_ptr, bflag, bond_ptr, (int)nbond, b_edge, b_i, s2)) != hipSuccess)
        return e;
    }
    side_what = "van der Waals";
    if ((e = xr_vdw(at, ff, row_ptr, nbr_j, nbr_s, acc, s2)) != hipSuccess) return e;
    side_what = "bond orders";
    XR_TIMED(XT_BOND_FORWARD, s2, e = xr_bond_forward(at, bd, ff, nbr_j, nbr_s, acc, s2));
    if (e != hipSuccess) return e;
    side_what = "many-body";
    XR_TIMED(XT_MANYBODY, s2,
             e = xr_manybody(at, bd, ff, row_ptr, hb_cand, hb_count, lammps, acc, tor_long, d_flags + 8, s2));
    if (e != hipSuccess) return e;
    side_what = "adjoint";
    XR_TIMED(XT_BACKWARD, s2, e = xr_bond_backward(at, bd, ff, s2));
    if (e != hipSuccess) return e;
    side_what = "event";
    return hipEventRecord(ctx->side_done, s2);
  };
  using SideLaunch = decltype(launch_side);
  bool side_issued = false;
  auto join_side = [&]() { return hipStreamWaitEvent(s, ctx->side_done, 0); };

  // ------------------------------------------------------------------ QEq
  CgArgs cg{};
  cg.S = S;
  cg.nd = nd;
  cg.nd_small = nd_small;
  cg.nu = nu;
  cg.type = in.types;
  cg.ff = ff;
  cg.row_ptr = row_ptr;
  cg.nbr_j = nbr_j;
  cg.col16 = col16;
  cg.hval = hval;
  cg.hval32 = hval32;
  {
    double2* vecs[10];
    for (int q = 0; q < 10; ++q) vecs[q] = cgv + q * (size_t)n;
    cg.v = CgVec{vecs[0], vecs[1], vecs[2], vecs[3], vecs[4], vecs[5], vecs[6], vecs[7],
                 {vecs[8], vecs[9]}, {cgq, cgq + n}};
  }
  cg.chunks = cg_chunks;
  cg.blocks = cg_blocks;
  cg.join = join;
  cg.minv = minv;
  cg.n = n;
  cg.binfo = binfo;
  cg.eta = eta;
  cg.partials = cg_part;
  cg.state = dc + L.cg;
  cg.tol = opt->qeq_tolerance > 0 ? opt->qeq_tolerance : 1e-10;
  cg.max_iter = opt->qeq_max_iterations > 0 ? opt->qeq_max_iterations : 2000;
  // Read back convergence first just past the launch the previous solve settled in, but no later than
  // twice the median of the recent solves (and never past 128 launches): one slow solve must not make
  // the next, easier one queue a long run of idle launches before its first read-back.
  {
    int recent[8], count = 0;
    for (int h : st.cg_history)
      if (h > 0) recent[count++] = h;
    int first = st.last_cg_launches > 0 ? st.last_cg_launches + 1 : 16;
    if (count > 0) {
      std::sort(recent, recent + count);
      first = std::min(first, 2 * recent[count / 2] + 1);
    }
    cg.first_check = std::max(2, std::min(first, 128));
  }
  cg.check_every = 2;
  cg.fixed_launches = cap ? cap->cg_launches : 0;
  cg.guess = (opt->warm_start && out.qeq_state) ? reinterpret_cast<const double2*>(out.qeq_state) : nullptr;
  cg.keep = reinterpret_cast<double2*>(out.qeq_state);
  cg.q = out.charges;
  cg.mu = reinterpret_cast<double*>(dc + L.mu);
  cg.abort_flag = d_flags;
  cg.poll_host = ctx->poll_host;
  cg.poll_ev[0] = ctx->poll_ev[0];
  cg.poll_ev[1] = ctx->poll_ev[1];
  hipError_t side_err = hipSuccess;
  struct SideHook {
    SideLaunch* fn;
    bool* issued;
    hipError_t* err;
  } hook{&launch_side, &side_issued, &side_err};
  cg.after_start = [](void* p) {
    SideHook* h = static_cast<SideHook*>(p);
    *h->issued = true;
    return *h->err = (*h->fn)();
  };
  cg.after_start_arg = &hook;
  // After the charges: Coulomb, forces, energies and the copies back.  With no
  // charge response to solve in between, this is issued behind the CG's first
  // chunk of launches when that chunk is expected to finish the solve (one
  // system, chunk sized from the previous solve), so the device goes on without
  // waiting for the host to see the convergence.  Every step overwrites its
  // outputs, so a speculation that the solve outran is simply issued again.
  const char* tail_what = "post-QEq";
  auto tail_front = [&]() -> hipError_t {
    hipError_t e = xr_nonbonded(at, ff, row_ptr, nbr_j, nbr_s, acc, dEdq, s);
    tail_what = "nonbonded";
    if (e == hipSuccess) e = join_side();
    if (e == hipSuccess) e = xr_force_assembly(at, bd, out.forces, s);
    if (e != hipSuccess) tail_what = "forces";
    return e;
  };
  auto tail_back = [&]() -> hipError_t {
    tail_what = "energy";
    hipError_t e = xr_energy_reduce(S, acc, reinterpret_cast<double*>(dc + L.comps), s);
    if (e == hipSuccess) e = hipMemcpyAsync(hc, dc, L.sys, hipMemcpyDeviceToHost, s);
    if (e != hipSuccess && tail_what[0] == 'e') tail_what = "copy";
    return e;
  };
  auto tail_all = [&]() -> hipError_t {
    hipError_t e = tail_front();
    return e == hipSuccess ? tail_back() : e;
  };
  using TailAll = decltype(tail_all);
  struct TailHook {
    TailAll* fn;
    hipError_t err;
  } tail{&tail_all, hipSuccess};
  if (!cap && !full && ns == 1 && st.last_cg_launches > 0) {
    cg.tail = [](void* p) {
      TailHook* t = static_cast<TailHook*>(p);
      return t->err = (*t->fn)();
    };
    cg.tail_arg = &tail;
  }
  bool tail_done = false;
  if (!given_q) {
    CgOutcome cgo{};
    const hipError_t e = xr_cg_solve(cg, s, &cgo);
    if (side_issued) HIPCHK(join_side(), "event");
    if (side_err != hipSuccess) return hip_fail(ctx, side_err, side_what);
    if (tail.err != hipSuccess) return hip_fail(ctx, tail.err, tail_what);
    HIPCHK(e, "QEq");
    if (cgo.aborted) {
      *overflow = true;
      return XREAC_OK;
    }
    tail_done = cgo.tail_done;
  } else {
    HIPCHK(launch_side(), side_what);
  }

  // ----------------------------------------------------- energy and forces
  if (!tail_done) {
    if (hipError_t e = tail_front(); e != hipSuccess) return hip_fail(ctx, e, tail_what);
    if (full) {
      // Charge response: neutral adjoint KKT solve with right-hand sides (dE/dq, -1).
      CgArgs adj = cg;
      adj.rhs = dEdq;
      adj.neutral = true;
      adj.guess = nullptr;
      adj.keep = nullptr;
      adj.q = lam;
      adj.mu = nullptr;
      adj.state = dc + L.cg2;
      adj.after_start = nullptr;
      adj.tail = nullptr;
      CgOutcome o2{};
      HIPCHK(xr_cg_solve(adj, s, &o2), "QEq adjoint");
      if (o2.aborted) {
        *overflow = true;
        return XREAC_OK;
      }
      HIPCHK(xr_charge_response(at, ff, row_ptr, nbr_j, nbr_s, lam, out.forces, s), "charge response");
    }
    if (hipError_t e = tail_back(); e != hipSuccess) return hip_fail(ctx, e, tail_what);
  }
  Finish fin;
  fin.L = L;
  fin.ns = ns;
  fin.n = n;
  fin.given_q = given_q;
  fin.full = full;
  fin.supplied = supplied;
  fin.fixed = cap != nullptr;
  fin.skin = skin;
  fin.m_cap = m_cap;
  fin.nbond = nbond;
  fin.nbr_count = in.nbr_count;
  fin.row_ptr = row_ptr;
  fin.nbr_j = nbr_j;
  fin.nbr_s = nbr_s;
  fin.bond_ptr = bond_ptr;
  fin.bond_i = b_i;
  fin.bond_j = bd.bj;
  fin.bond_code = bd.code;
  fin.bond_BO = bd.BO;
  if (cap) {  // recorded, not run: the replay's host side reads the control block
    cap->fin = fin;
    return XREAC_OK;
  }
  HIPCHK(hipStreamSynchronize(s), "final sync");
  if (!arena) ctx->prof.resolve();
  return finish_eval(ctx, st, fin, out, overflow);
}

int run_batch(xreac_ctx* ctx, const xreac_ff* ffh, const BatchIn& in, const xreac_options* opt, BatchOut& out,
              Arena* arena = nullptr) {
  // The CG sweep addresses gathered vector entries with 32-bit byte offsets (16 B per atom).
  if (in.offsets[in.ns] >= (1 << 28)) return fail(ctx, XREAC_INVALID, "at most 268,435,455 atoms per call");
  std::lock_guard<std::mutex> serial(g_eval_lock[ctx->device]);
  bool overflow = false;
  // With a warm start the QEq of the first attempt writes its solution over
  // the guess (result->qeq_state).  An attempt whose buffers (sized from the
  // last evaluation) overflowed is repeated from the guess as it was, so a
  // result never depends on whether the sizes of the last call fitted.
  const size_t guess_bytes = sizeof(double2) * (size_t)in.offsets[in.ns];
  const bool keep_guess = !arena && opt->warm_start && out.qeq_state;
  if (keep_guess) {
    HIPCHK(hipSetDevice(ctx->device), "hipSetDevice");
    HIPCHK(ctx->es.guess_buf.ensure(guess_bytes), "guess copy");
    HIPCHK(hipMemcpyAsync(ctx->es.guess_buf.ptr, out.qeq_state, guess_bytes, hipMemcpyDeviceToDevice, ctx->stream),
           "guess copy");
  }
  int rc = evaluate_once(ctx, ctx->es, ffh, in, opt, out, false, &overflow, arena);
  if (rc == XREAC_OK && overflow) {
    // The overflowed attempt can still have work in flight (launches queued
    // past the read-back that saw the overflow, and the side stream's bonded
    // terms, which write into the bond workspace); the repeat may free and
    // regrow those buffers, and freed memory can go to another context at
    // once, so nothing of the attempt may run after this point.
    HIPCHK(hipStreamSynchronize(ctx->side), "overflow sync");
    HIPCHK(hipStreamSynchronize(ctx->stream), "overflow sync");
    if (keep_guess)
      HIPCHK(hipMemcpyAsync(out.qeq_state, ctx->es.guess_buf.ptr, guess_bytes, hipMemcpyDeviceToDevice, ctx->stream),
             "guess copy");
    rc = evaluate_once(ctx, ctx->es, ffh, in, opt, out, true, &overflow, arena);
  }
  if (rc == XREAC_OK && overflow) return fail(ctx, XREAC_DEVICE, "internal error: buffer sizing");
  return rc;
}

}  // namespace

extern "C" {

}  // extern "C"

namespace {

int evaluate_single(xreac_ctx* ctx, const xreac_ff* ffh, const xreac_system* sys, const xreac_options* opt,
                    xreac_result* out, Arena* arena) {
  if (!ctx || !ffh || !sys || !opt || !out) return XREAC_INVALID;
  ctx->message.clear();
  if (sys->n <= 0 || !sys->positions || !sys->types || !out->forces || !out->charges)
    return fail(ctx, XREAC_INVALID, "A nonempty system with device positions, types, forces and charges is required");
  const int32_t offsets[2] = {0, sys->n};
  BatchIn in{1, offsets, sys->positions, sys->types, sys->cell, sys->pbc, &sys->total_charge};
  in.nbr_row_ptr = sys->neighbor_row_ptr;
  in.nbr_j = sys->neighbor_j;
  in.nbr_shifts = sys->neighbor_shifts;
  in.nbr_count = sys->neighbor_count;
  if (in.nbr_row_ptr && (!in.nbr_j || !in.nbr_shifts || in.nbr_count < 0 || in.nbr_count > INT32_MAX - 1))
    return fail(ctx, XREAC_INVALID, "a supplied neighbor list needs row pointers, j, shifts and a count");
  int32_t status = XREAC_OK, iters = 0;
  BatchOut bo{};
  bo.comps = out->components;
  bo.status = &status;
  bo.iters = &iters;
  bo.forces = out->forces;
  bo.charges = out->charges;
  bo.tbo = out->total_bond_orders;
  bo.nlp = out->lone_pairs;
  bo.bcounts = out->bond_counts;
  bo.qeq_state = nullptr;
  xreac_options o = *opt;
  o.warm_start = 0;
  const int rc = run_batch(ctx, ffh, in, &o, bo, arena);
  if (rc != XREAC_OK) return rc;
  out->num_edges = bo.num_edges;
  out->num_bonds = bo.num_bonds;
  out->qeq_iterations = iters;
  if (status != XREAC_OK) return fail(ctx, status, ctx->sys_messages[0]);
  return XREAC_OK;
}

}  // namespace

extern "C" {

int xreac_evaluate(xreac_ctx* ctx, const xreac_ff* ffh, const xreac_system* sys, const xreac_options* opt,
                   xreac_result* out) {
  return evaluate_single(ctx, ffh, sys, opt, out, nullptr);
}

int xreac_evaluate_static(xreac_ctx* ctx, const xreac_ff* ffh, const xreac_system* sys, const xreac_options* opt,
                          xreac_result* out) {
  if (!ctx) return XREAC_INVALID;
  if (!ctx->scratch) return fail(ctx, XREAC_DEVICE, "xreac_evaluate_static: no static scratch on this device");
  if (opt && opt->neighbor_skin != 0.0)
    return fail(ctx, XREAC_INVALID, "xreac_evaluate_static keeps no skin lists: neighbor_skin must be 0");
  Arena arena;
  arena.base = static_cast<char*>(ctx->scratch);
  arena.cap = XREAC_STATIC_SCRATCH_BYTES;
  return evaluate_single(ctx, ffh, sys, opt, out, &arena);
}

int xreac_evaluate_batch(xreac_ctx* ctx, const xreac_ff* ffh, const xreac_batch* b, const xreac_options* opt,
                         xreac_batch_result* out) {
  if (!ctx || !ffh || !b || !opt || !out) return XREAC_INVALID;
  ctx->message.clear();
  if (b->num_systems <= 0 || !b->atom_offsets || !b->positions || !b->types || !out->forces || !out->charges ||
      !out->components || !out->status || !out->qeq_iterations)
    return fail(ctx, XREAC_INVALID, "A nonempty batch with offsets, device positions, types, forces and charges "
                                    "and host component/status/iteration arrays is required");
  if (b->atom_offsets[0] != 0) return fail(ctx, XREAC_INVALID, "atom_offsets[0] must be 0");
  for (int q = 0; q < b->num_systems; ++q)
    if (b->atom_offsets[q + 1] <= b->atom_offsets[q])
      return fail(ctx, XREAC_INVALID, "every system of a batch needs at least one atom");
  if (opt->warm_start && !out->qeq_state) return fail(ctx, XREAC_INVALID, "warm_start needs result->qeq_state");
  BatchIn in{b->num_systems, b->atom_offsets, b->positions, b->types, b->cells, b->pbc, b->total_charges};
  in.nbr_row_ptr = b->neighbor_row_ptr;
  in.nbr_j = b->neighbor_j;
  in.nbr_shifts = b->neighbor_shifts;
  in.nbr_count = b->neighbor_count;
  if (in.nbr_row_ptr && (!in.nbr_j || !in.nbr_shifts || in.nbr_count < 0 || in.nbr_count > INT32_MAX - 1))
    return fail(ctx, XREAC_INVALID, "a supplied neighbor list needs row pointers, j, shifts and a count");
  BatchOut bo{};
  bo.comps = out->components;
  bo.status = out->status;
  bo.iters = out->qeq_iterations;
  bo.forces = out->forces;
  bo.charges = out->charges;
  bo.tbo = out->total_bond_orders;
  bo.nlp = out->lone_pairs;
  bo.bcounts = out->bond_counts;
  bo.qeq_state = out->qeq_state;
  const int rc = run_batch(ctx, ffh, in, opt, bo);
  if (rc != XREAC_OK) return rc;
  out->num_edges = bo.num_edges;
  out->num_bonds = bo.num_bonds;
  return XREAC_OK;
}

int xreac_verlet(xreac_ctx* ctx, const xreac_ff* ffh, int32_t ns, const int32_t* offsets, const int32_t* types,
                 const double* dt_fs, double* positions, double* velocities, const double* forces, int32_t stage,
                 double* kinetic) {
  if (!ctx || !ffh || ns <= 0 || !offsets || !types || !dt_fs || !positions || !velocities || !forces ||
      (stage != 0 && stage != 1))
    return XREAC_INVALID;
  ctx->message.clear();
  HIPCHK(hipSetDevice(ctx->device), "hipSetDevice");
  HIPCHK(xr_verlet(ffh->dev, ns, offsets, types, dt_fs, positions, velocities, forces, stage, kinetic, ctx->stream),
         "verlet");
  return XREAC_OK;
}

int xreac_qeq_extrapolate(xreac_ctx* ctx, int32_t ns, int32_t num_atoms, const int32_t* offsets,
                          const int32_t* order, double* qeq_state, double* history) {
  if (!ctx || ns <= 0 || num_atoms < 0 || !offsets || !order || !qeq_state || !history) return XREAC_INVALID;
  ctx->message.clear();
  HIPCHK(hipSetDevice(ctx->device), "hipSetDevice");
  HIPCHK(xr_qeq_extrapolate(ns, num_atoms, offsets, order, qeq_state, history, ctx->stream), "QEq extrapolation");
  return XREAC_OK;
}

int xreac_fire_step(xreac_ctx* ctx, int32_t ns, const int32_t* offsets, double* positions, double* velocities,
                    const double* forces, double* state, int32_t* status, double force_tolerance,
                    int32_t max_iterations) {
  if (!ctx || ns <= 0 || !offsets || !positions || !velocities || !forces || !state || !status ||
      !(force_tolerance > 0) || max_iterations < 0)
    return XREAC_INVALID;
  ctx->message.clear();
  HIPCHK(hipSetDevice(ctx->device), "hipSetDevice");
  HIPCHK(xr_fire(ns, offsets, positions, velocities, forces, state, status, force_tolerance, max_iterations,
                 ctx->stream),
         "FIRE");
  return XREAC_OK;
}

int64_t xreac_neighbor_list_builds(const xreac_ctx* ctx) { return ctx ? ctx->neighbor_list_builds : -1; }

const char* xreac_batch_error(const xreac_ctx* ctx, int32_t system) {
  if (!ctx || system < 0 || system >= (int32_t)ctx->sys_messages.size()) return "no such system";
  return ctx->sys_messages[system].c_str();
}

int xreac_profile_enable(xreac_ctx* ctx, int timer_mask) {
  if (!ctx) return XREAC_INVALID;
  ctx->prof.mask = (unsigned)timer_mask & ((1u << XT_COUNT) - 1u);
  ctx->prof.on = ctx->prof.mask != 0;
  return XREAC_OK;
}

int xreac_profile_stride(xreac_ctx* ctx, int every) {
  if (!ctx || every < 1) return XREAC_INVALID;
  ctx->prof.every = every;
  for (long long& c : ctx->prof.seen) c = 0;
  return XREAC_OK;
}

int xreac_profile_read(xreac_ctx* ctx, int timer, double* total_ms, int64_t* launches, int reset) {
  if (!ctx || timer < 0 || timer >= XT_COUNT || !total_ms || !launches) return XREAC_INVALID;
  *total_ms = ctx->prof.total_ms[timer];
  *launches = ctx->prof.launches[timer];
  if (reset) {
    ctx->prof.total_ms[timer] = 0.0;
    ctx->prof.launches[timer] = 0;
  }
  return XREAC_OK;
}

int xreac_copy_neighbors(xreac_ctx* ctx, int32_t* row_ptr, int32_t* j, int32_t* shifts, int64_t capacity) {
  if (!ctx || !row_ptr || !j || !shifts) return XREAC_INVALID;
  if (capacity < ctx->es.m) return fail(ctx, XREAC_INVALID, "neighbor capacity too small");
  HIPCHK(hipSetDevice(ctx->device), "hipSetDevice");
  std::vector<int> codes(ctx->es.m);
  HIPCHK(hipMemcpy(row_ptr, ctx->es.row_ptr, sizeof(int) * (ctx->es.n + 1), hipMemcpyDeviceToHost), "copy");
  if (ctx->es.m) {
    HIPCHK(hipMemcpy(j, ctx->es.nbr_j, sizeof(int) * ctx->es.m, hipMemcpyDeviceToHost), "copy");
    HIPCHK(hipMemcpy(codes.data(), ctx->es.nbr_s, sizeof(int) * ctx->es.m, hipMemcpyDeviceToHost), "copy");
  }
  for (int64_t e = 0; e < ctx->es.m; ++e) {
    int sx, sy, sz;
    xr_unpack_shift(codes[e], sx, sy, sz);
    shifts[3 * e] = sx;
    shifts[3 * e + 1] = sy;
    shifts[3 * e + 2] = sz;
  }
  return XREAC_OK;
}

int xreac_copy_bonds(xreac_ctx* ctx, int32_t* i, int32_t* j, int32_t* shifts, double* bo, int64_t capacity) {
  if (!ctx || !i || !j || !shifts || !bo) return XREAC_INVALID;
  if (capacity < ctx->es.nb) return fail(ctx, XREAC_INVALID, "bond capacity too small");
  HIPCHK(hipSetDevice(ctx->device), "hipSetDevice");
  std::vector<int> codes(ctx->es.nb);
  if (ctx->es.nb) {
    HIPCHK(hipMemcpy(i, ctx->es.bond_i, sizeof(int) * ctx->es.nb, hipMemcpyDeviceToHost), "copy");
    HIPCHK(hipMemcpy(j, ctx->es.bond_j, sizeof(int) * ctx->es.nb, hipMemcpyDeviceToHost), "copy");
    HIPCHK(hipMemcpy(codes.data(), ctx->es.bond_code, sizeof(int) * ctx->es.nb, hipMemcpyDeviceToHost), "copy");
    HIPCHK(hipMemcpy(bo, ctx->es.bond_BO, sizeof(double) * ctx->es.nb, hipMemcpyDeviceToHost), "copy");
  }
  for (int64_t e = 0; e < ctx->es.nb; ++e) {
    int sx, sy, sz;
    xr_unpack_shift(codes[e], sx, sy, sz);
    shifts[3 * e] = sx;
    shifts[3 * e + 1] = sy;
    shifts[3 * e + 2] = sz;
  }
  return XREAC_OK;
}

}  // extern "C"

// ------------------------------------------------------------ recorded steps
// One step of a fixed batch (integrator stages around an evaluation) recorded
// once as a HIP graph and replayed, so a step costs one graph launch and one
// wait instead of a hundred host-issued launches and the QEq read-backs.  The
// step owns its workspaces (their addresses are baked into the graph); the
// context's own are untouched by replays.  A replay whose buffers overflowed
// or whose QEq needed more launches than were recorded restores what the
// step modifies and runs the step again with the context's read-backs, then
// records it afresh, so a run always completes the step with the results the
// unrecorded step gives.
struct xreac_step {
  xreac_ctx* ctx = nullptr;
  int device = 0;
  const xreac_ff* ff = nullptr;
  EvalState st;
  int ns = 0, n = 0;
  std::vector<int32_t> offsets, pbc;
  std::vector<double> cells, charges;
  const double* x = nullptr;
  const int32_t* types = nullptr;
  xreac_options opt{};
  xreac_step_stages stg{};
  double* forces = nullptr;
  double* q = nullptr;
  double* tbo = nullptr;
  double* nlp = nullptr;
  int32_t* bcounts = nullptr;
  double* qeq_state = nullptr;
  Buffer dev_offsets;  // device copy of offsets (integrator stages)
  Buffer backup;       // what the step modifies, as it was before the replay
  struct Saved {
    void* at;
    size_t bytes;
    size_t off;
  };
  std::vector<Saved> saved;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  Finish fin;
  int cg_launches = 0;
  int need_max = 0;      // most QEq launches a replay needed since the last recording
  int since_record = 0;  // replays since the last recording
  int64_t replays = 0, fallbacks = 0, recordings = 0;
  ~xreac_step() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
  }
};

namespace {

BatchIn step_batch(const xreac_step* p) {
  BatchIn in{p->ns, p->offsets.data(), p->x, p->types, p->cells.empty() ? nullptr : p->cells.data(),
             p->pbc.empty() ? nullptr : p->pbc.data(), p->charges.empty() ? nullptr : p->charges.data()};
  return in;
}

BatchOut step_out(xreac_step* p, xreac_batch_result* r) {
  BatchOut bo{};
  bo.comps = r->components;
  bo.status = r->status;
  bo.iters = r->qeq_iterations;
  bo.forces = p->forces;
  bo.charges = p->q;
  bo.tbo = p->tbo;
  bo.nlp = p->nlp;
  bo.bcounts = p->bcounts;
  bo.qeq_state = p->qeq_state;
  return bo;
}

// The arrays a step modifies, to (save) or from (restore) the step's backup, in one launch.
hipError_t step_backup(xreac_step* p, bool save, hipStream_t s) {
  CopyRanges c{};
  for (const auto& v : p->saved) {
    auto* b = reinterpret_cast<unsigned*>(static_cast<char*>(p->backup.ptr) + v.off);
    auto* a = static_cast<unsigned*>(v.at);
    c.src[c.count] = save ? a : b;
    c.dst[c.count] = save ? b : a;
    c.words[c.count] = (long long)(v.bytes / 4);
    ++c.count;
  }
  if (c.count == 0) return hipSuccess;
  hipLaunchKernelGGL(k_copy_ranges, dim3(256), dim3(256), 0, s, c);
  return hipGetLastError();
}

hipError_t stage_before(xreac_step* p, hipStream_t s) {
  const int* off = static_cast<const int*>(p->dev_offsets.ptr);
  hipError_t e = hipSuccess;
  if (p->stg.dt_fs)
    e = xr_verlet(p->ff->dev, p->ns, off, p->types, p->stg.dt_fs, const_cast<double*>(p->x), p->stg.velocities,
                  p->forces, 0, nullptr, s);
  if (e == hipSuccess && p->stg.qeq_order)
    e = xr_qeq_extrapolate(p->ns, p->n, off, p->stg.qeq_order, p->qeq_state, p->stg.qeq_history, s);
  return e;
}

hipError_t stage_after(xreac_step* p, hipStream_t s) {
  const int* off = static_cast<const int*>(p->dev_offsets.ptr);
  hipError_t e = hipSuccess;
  if (p->stg.dt_fs)
    e = xr_verlet(p->ff->dev, p->ns, off, p->types, p->stg.dt_fs, const_cast<double*>(p->x), p->stg.velocities,
                  p->forces, 1, p->stg.kinetic, s);
  if (e == hipSuccess && p->stg.fire_state)
    e = xr_fire(p->ns, off, const_cast<double*>(p->x), p->stg.velocities, p->forces, p->stg.fire_state,
                p->stg.fire_status, p->stg.fire_force_tolerance, p->stg.fire_max_iterations, s);
  return e;
}

// The step with the context's read-backs (the first run, and a replay that
// did not complete).
int step_unrecorded(xreac_step* p, xreac_batch_result* r) {
  xreac_ctx* ctx = p->ctx;
  HIPCHK(stage_before(p, ctx->stream), "step stages");
  BatchIn in = step_batch(p);
  BatchOut bo = step_out(p, r);
  const int rc = run_batch(ctx, p->ff, in, &p->opt, bo);
  if (rc != XREAC_OK) return rc;
  r->num_edges = bo.num_edges;
  r->num_bonds = bo.num_bonds;
  HIPCHK(stage_after(p, ctx->stream), "step stages");
  HIPCHK(hipStreamSynchronize(ctx->stream), "step sync");
  return XREAC_OK;
}

// QEq launches of a recording that most recently needed `need`: an eighth
// more plus three.  Every launch past the last system's settling is an idle
// dispatch in each replay; a replay that needs more falls back and records
// again.
int recorded_launches(int need) { return std::min(need + need / 8 + 3, 4096); }

// Record the step, sized from the context's latest sizing history (after an
// unrecorded run on it) or from the step's own (replays since the last
// recording; `need` then gives the QEq launches).  A workspace that would
// have to grow while the stream records is grown first and the recording
// repeated (each buffer grows at most once per recording).
int step_record(xreac_step* p, int need = -1) {
  xreac_ctx* ctx = p->ctx;
  hipStream_t s = ctx->stream;
  if (p->exec) (void)hipGraphExecDestroy(p->exec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  p->exec = nullptr;
  p->graph = nullptr;
  if (need < 0) {
    // from the unrecorded run just made (not the most of recent solves: one
    // cold start before a run of warm ones would set the count for all)
    p->st.take_hints(ctx->es);
    need = p->st.last_cg_launches;
  }
  p->cg_launches = recorded_launches(need);
  p->need_max = 0;
  p->since_record = 0;
  size_t backup_bytes = 0;
  for (const auto& v : p->saved) backup_bytes = std::max(backup_bytes, v.off + v.bytes);
  HIPCHK(p->backup.ensure(backup_bytes + 256), "step backup");
  BatchIn in = step_batch(p);
  std::vector<double> comps((size_t)p->ns * NCOMP);
  std::vector<int32_t> status(p->ns), iters(p->ns);
  xreac_batch_result r{};
  r.components = comps.data();
  r.status = status.data();
  r.qeq_iterations = iters.data();
  BatchOut bo = step_out(p, &r);
  for (int attempt = 0; attempt < 16; ++attempt) {
    Capture cap;
    cap.cg_launches = p->cg_launches;
    HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "step recording");
    hipError_t e = step_backup(p, true, s);
    if (e == hipSuccess) e = stage_before(p, s);
    bool overflow = false;
    int rc = XREAC_OK;
    if (e == hipSuccess) rc = evaluate_once(ctx, p->st, p->ff, in, &p->opt, bo, false, &overflow, nullptr, &cap);
    const bool complete = e == hipSuccess && rc == XREAC_OK && !cap.grow && !cap.grow_pinned;
    if (complete) e = stage_after(p, s);
    hipGraph_t g = nullptr;
    const hipError_t ee = hipStreamEndCapture(s, &g);
    if (complete && e == hipSuccess && ee == hipSuccess) {
      p->graph = g;
      const hipError_t ei = hipGraphInstantiate(&p->exec, g, nullptr, nullptr, 0);
      if (ei != hipSuccess) return hip_fail(ctx, ei, "step graph instantiation");
      p->fin = cap.fin;
      ++p->recordings;
      return XREAC_OK;
    }
    if (g) (void)hipGraphDestroy(g);
    if (rc != XREAC_OK) return rc;
    if (e != hipSuccess) return hip_fail(ctx, e, "step recording");
    if (ee != hipSuccess) return hip_fail(ctx, ee, "step recording");
    if (cap.grow) HIPCHK(cap.grow->ensure(cap.grow_bytes), "step workspace");
    if (cap.grow_pinned) {
      if (p->st.ctl_host) (void)hipHostFree(p->st.ctl_host);
      p->st.ctl_host = nullptr;
      p->st.ctl_host_bytes = 0;
      HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&p->st.ctl_host), cap.grow_pinned), "pinned control block");
      p->st.ctl_host_bytes = cap.grow_pinned;
    }
  }
  return fail(ctx, XREAC_DEVICE, "internal error: step workspaces kept growing while recording");
}

}  // namespace

extern "C" {

int xreac_step_create(xreac_ctx* ctx, const xreac_ff* ffh, const xreac_batch* b, const xreac_options* opt,
                      const xreac_step_stages* stages, const xreac_batch_result* result, xreac_step** out) {
  if (!ctx || !ffh || !b || !opt || !result || !out) return XREAC_INVALID;
  ctx->message.clear();
  *out = nullptr;
  if (b->num_systems <= 0 || !b->atom_offsets || !b->positions || !b->types || !result->forces || !result->charges)
    return fail(ctx, XREAC_INVALID, "A nonempty batch with offsets, device positions, types, forces and charges is "
                                    "required");
  if (b->atom_offsets[0] != 0) return fail(ctx, XREAC_INVALID, "atom_offsets[0] must be 0");
  for (int q = 0; q < b->num_systems; ++q)
    if (b->atom_offsets[q + 1] <= b->atom_offsets[q])
      return fail(ctx, XREAC_INVALID, "every system of a batch needs at least one atom");
  if (b->atom_offsets[b->num_systems] >= (1 << 28)) return fail(ctx, XREAC_INVALID, "at most 268,435,455 atoms per call");
  if (b->neighbor_row_ptr) return fail(ctx, XREAC_INVALID, "a recorded step builds its neighbor lists on the device");
  if (opt->neighbor_skin != 0.0)
    return fail(ctx, XREAC_INVALID, "a recorded step builds fresh neighbor lists: neighbor_skin must be 0");
  if (opt->warm_start && !result->qeq_state) return fail(ctx, XREAC_INVALID, "warm_start needs result->qeq_state");
  if (!ctx->stream) return fail(ctx, XREAC_INVALID, "a recorded step needs a context created with its own stream");
  xreac_step_stages none{};
  const xreac_step_stages& sg = stages ? *stages : none;
  if (sg.dt_fs && !sg.velocities) return fail(ctx, XREAC_INVALID, "velocity Verlet stages need velocities");
  if (sg.qeq_order && (!sg.qeq_history || !result->qeq_state))
    return fail(ctx, XREAC_INVALID, "QEq extrapolation needs qeq_history and result->qeq_state");
  if (sg.fire_state && (!sg.velocities || !sg.fire_status || !(sg.fire_force_tolerance > 0) ||
                        sg.fire_max_iterations < 0))
    return fail(ctx, XREAC_INVALID, "a FIRE stage needs velocities, status, a positive tolerance and a limit");
  auto* p = new xreac_step();
  p->ctx = ctx;
  p->device = ctx->device;
  p->ff = ffh;
  p->ns = b->num_systems;
  p->n = b->atom_offsets[b->num_systems];
  p->offsets.assign(b->atom_offsets, b->atom_offsets + p->ns + 1);
  if (b->cells) p->cells.assign(b->cells, b->cells + 9 * (size_t)p->ns);
  if (b->pbc) p->pbc.assign(b->pbc, b->pbc + 3 * (size_t)p->ns);
  if (b->total_charges) p->charges.assign(b->total_charges, b->total_charges + p->ns);
  p->x = b->positions;
  p->types = b->types;
  p->opt = *opt;
  p->stg = sg;
  p->forces = result->forces;
  p->q = result->charges;
  p->tbo = result->total_bond_orders;
  p->nlp = result->lone_pairs;
  p->bcounts = result->bond_counts;
  p->qeq_state = result->qeq_state;
  const size_t n = (size_t)p->n, ns = (size_t)p->ns;
  size_t off = 0;
  auto save = [&](const void* at, size_t bytes) {
    if (!at) return;
    p->saved.push_back({const_cast<void*>(at), bytes, off});
    off = (off + bytes + 255) & ~size_t(255);
  };
  const bool moves = sg.dt_fs || sg.fire_state;
  if (moves) save(p->x, 24 * n);
  if (moves) save(sg.velocities, 24 * n);
  if (sg.dt_fs) save(p->forces, 24 * n);  // stage 0 reads the previous step's forces
  if (opt->warm_start || sg.qeq_order) save(p->qeq_state, 16 * n);
  if (sg.qeq_order) save(sg.qeq_history, 48 * n);
  if (sg.fire_state) save(sg.fire_state, 32 * ns);
  if (sg.fire_state) save(sg.fire_status, 4 * ns);
  hipError_t e = hipSetDevice(ctx->device);
  if (e == hipSuccess) e = p->dev_offsets.ensure(sizeof(int32_t) * (ns + 1));
  if (e == hipSuccess)
    e = hipMemcpy(p->dev_offsets.ptr, b->atom_offsets, sizeof(int32_t) * (ns + 1), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    delete p;
    return hip_fail(ctx, e, "step set-up");
  }
  *out = p;
  return XREAC_OK;
}

int xreac_step_run(xreac_step* p, xreac_batch_result* r) {
  if (!p || !r || !r->components || !r->status || !r->qeq_iterations) return XREAC_INVALID;
  xreac_ctx* ctx = p->ctx;
  ctx->message.clear();
  HIPCHK(hipSetDevice(ctx->device), "hipSetDevice");
  if (p->exec) {
    BatchOut bo = step_out(p, r);
    bool incomplete = false;
    {
      std::lock_guard<std::mutex> serial(g_eval_lock[ctx->device]);
      HIPCHK(hipGraphLaunch(p->exec, ctx->stream), "step replay");
      HIPCHK(hipStreamSynchronize(ctx->stream), "step replay");
    }
    const int rc = finish_eval(ctx, p->st, p->fin, bo, &incomplete);
    if (rc != XREAC_OK) return rc;
    if (!incomplete) {
      ++p->replays;
      r->num_edges = bo.num_edges;
      r->num_bonds = bo.num_bonds;
      // Eight replays that all settled well inside the recorded launch count
      // (the recording run was the slowest, e.g. the first warm step of a
      // segment): record again with fewer launches.
      p->need_max = std::max(p->need_max, p->st.last_cg_launches);
      if (++p->since_record >= 8) {
        if (recorded_launches(p->need_max) < p->cg_launches) return step_record(p, p->need_max);
        p->need_max = 0;
        p->since_record = 0;
      }
      return XREAC_OK;
    }
    // put back what the replay changed, then run the step with read-backs
    HIPCHK(step_backup(p, false, ctx->stream), "step restore");
    ++p->fallbacks;
  }
  const int rc = step_unrecorded(p, r);
  if (rc != XREAC_OK) return rc;
  return step_record(p);
}

int xreac_step_stats(const xreac_step* p, int64_t* replays, int64_t* fallbacks, int64_t* recordings) {
  if (!p) return XREAC_INVALID;
  if (replays) *replays = p->replays;
  if (fallbacks) *fallbacks = p->fallbacks;
  if (recordings) *recordings = p->recordings;
  return XREAC_OK;
}

void xreac_step_destroy(xreac_step* p) {
  if (!p) return;
  // Every run waits for its work, so nothing of the step is in flight here;
  // the context is not touched (it may already be gone).
  (void)hipSetDevice(p->device);
  delete p;
}

}  // extern "C"
