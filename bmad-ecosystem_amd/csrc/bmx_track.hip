// bmx_track.hip -- gfx950 tracking kernels and the libbmx C ABI.
//
// One particle per lane.  Each lane keeps its 6-D coordinates, beta and two
// per-particle invariants of the drift map in registers for the whole launch
// (all turns, all elements): the structure-of-arrays store in HBM is read once
// and written once per launch.  The element program (bmx_internal.h) is walked
// with a wave-uniform index, so op constants come through the scalar cache and
// the op dispatch is a scalar branch.  Lost lanes leave the loop early and
// record Bmad's loss state, element and turn.
//
// The arithmetic of every op restates the reference routine cited at it in
// FP64, with particle-independent factors hoisted by the lattice compiler and
// divisions shared where the algebra allows (1/(1+pz) and beta/beta0 - 1 are
// per-particle constants between energy kicks).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <thread>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <functional>
#include <vector>

#include "../../include/bmx.h"
#include "bmx_internal.h"
#include "bmx_jit.h"
#include "bmx_tangent.h"
#include <tuple>

namespace bmx {
bool compile_variant(const bmx_ele_rec* eles, int n_ele, const bmx_ref& ref, const bmx_config& cfg,
                     std::vector<DevOp>& prog, std::vector<int32_t>& ele_op_begin, LatticeConsts* lc,
                     std::string* err, const RadInput* rad = nullptr, std::vector<double>* rad_table = nullptr);
}

using namespace bmx;

#include "bmx_ops.h"
#include "bmx_kernels.h"

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
struct bmx_ctx {
  int device;
  hipStream_t stream;
  hipEvent_t ev0, ev1;
  std::string err;
  bmx_schedule sched;  // turn-block schedule used when a call passes turns_per_block <= 0
  void* scratch = nullptr;  // device address of this device's `scratch` array (bmx_track_bunch_noalloc)
  std::string scratch_err;
};

// ---------------------------------------------------------------------------
// Static scratch of bmx_track_bunch_noalloc: the lost-particle counter and
// staging copies of host-resident particle arrays, in one file-scope device
// array, so that the entry point allocates nothing.  Every GPU holds its own
// copy until the process ends; each context looks its address up once, at
// creation, with its device current.  Layout: the counter in the first 256
// bytes, then NOALLOC_MAX_PARTICLES slots of 8 double arrays and 3 int32
// arrays (x, px, y, py, z, pz, beta, one spare; state, ix_lost, turn_lost).
// ---------------------------------------------------------------------------
constexpr int64_t NOALLOC_MAX_PARTICLES = BMX_NOALLOC_MAX_PARTICLES;
constexpr size_t NOALLOC_HEAD = 256;
constexpr size_t NOALLOC_BYTES =
    NOALLOC_HEAD + (size_t)NOALLOC_MAX_PARTICLES * (8 * sizeof(double) + 3 * sizeof(int32_t));
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

static std::mutex g_scratch_mu;
static std::vector<void*> g_scratch_ptr;  // per device, filled once

// The device address of `scratch` on `device` (which must be current); null
// with the HIP error text in *err when the lookup fails.
static void* scratch_address(int device, std::string* err) {
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if ((int)g_scratch_ptr.size() <= device) g_scratch_ptr.resize(device + 1, nullptr);
  if (!g_scratch_ptr[device]) {
    void* ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      *err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
      return nullptr;
    }
    g_scratch_ptr[device] = ptr;
  }
  return g_scratch_ptr[device];
}

__global__ void bmx_zero_u64_kernel(unsigned long long* p) { *p = 0ull; }

struct bmx_lattice {
  int device;
  int n_variants;
  std::vector<int32_t> var_op_begin;            // n_variants + 1
  std::vector<std::vector<int32_t>> ele_op_begin;  // per variant, n_ele + 1 (absolute op indices)
  std::vector<int32_t> var_n_ele;
  DevOp* d_ops;
  int32_t* d_var_op_begin;
  int32_t* d_var_n_ele;
  int32_t* d_cell_prv;        // lc.cell_prv (cell kernels only)
  const JitKernels* jit;      // runtime-compiled cell kernels, or null: the interpreter kernels
  int32_t cell_len;
  int64_t n_cells;
  int64_t n_ops;
  LatticeConsts lc;
  // backward tracking of variant 0 (bmx_track_bunch with direction = -1):
  // its elements compiled at upload into a program of their own, so that no
  // tracking call compiles or allocates (d_ops_bwd stays null, with the
  // reason in bwd_err, when the backward path refuses the lattice)
  DevOp* d_ops_bwd = nullptr;
  std::vector<int32_t> eb_bwd;
  LatticeConsts lc_bwd;
  std::string bwd_err;
  // reference times of variant 0's element boundaries (LatticeConsts::rt), host and device
  std::vector<double> rt;
  double* d_rt = nullptr;
  bool abs_time = false;  // uploaded with absolute-time rf
  // radiation maps of every variant's OP_RAD ops (LatticeConsts::rad) and the
  // fluctuation stream: the number of the next call that draws deviates
  double* d_rad = nullptr;
  bool rad_fluct = false;
  mutable std::atomic<uint64_t> rad_next_call{0};
  // per variant: the lowest rf frequency among the cavities that are on (0:
  // none), which sets the rf wavelength of the closed-orbit search
  std::vector<double> var_rf_freq;
};

// The fluctuation stream of one tracking call: the lattice's next call number.
static void rad_call_setup(const bmx_lattice* lat, LatticeConsts& lc) {
  if (lat->rad_fluct) lc.rad_call = (uint32_t)lat->rad_next_call.fetch_add(1);
}

// Launch a tracking kernel: the lattice's runtime-compiled cell kernel when
// it has one, else the built-in one (same parameters).
template <class... KA, class... Args>
static void launch_track(hipFunction_t jf, void (*kern)(KA...), dim3 grid, dim3 block, hipStream_t st,
                         Args&&... args) {
  if (jf) {
    std::tuple<KA...> a(std::forward<Args>(args)...);
    void* p[sizeof...(KA)];
    std::apply([&](auto&... x) {
      int i = 0;
      ((p[i++] = (void*)&x), ...);
    }, a);
    (void)hipModuleLaunchKernel(jf, grid.x, grid.y, grid.z, block.x, block.y, block.z, 0, st, p, nullptr);
  } else {
    hipLaunchKernelGGL(kern, grid, block, 0, st, std::forward<Args>(args)...);
  }
}

static int set_err(bmx_ctx* ctx, const std::string& s) {
  if (ctx) ctx->err = s;
  return -1;
}

#define HIP_TRY(ctx, call)                                                              \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess) return set_err((ctx), std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

// Shared with bmx_mlp.hip (hidden: not part of the exported ABI).
extern "C" __attribute__((visibility("hidden"))) hipStream_t bmx_ctx_stream(const bmx_ctx* ctx) { return ctx->stream; }
extern "C" __attribute__((visibility("hidden"))) int bmx_ctx_device(const bmx_ctx* ctx) { return ctx->device; }
// defined in bmx_mlp.hip: frees the context's GEMM workspace
extern "C" void bmx_mlp_release(const bmx_ctx* ctx);
extern "C" __attribute__((visibility("hidden"))) int32_t bmx_ctx_set_error(bmx_ctx* ctx, const char* msg) {
  return set_err(ctx, msg ? msg : "error");
}

extern "C" {

int32_t bmx_abi_version(void) { return BMX_ABI_VERSION; }

int32_t bmx_sizeof(int32_t which) {
  switch (which) {
    case 0: return (int32_t)sizeof(bmx_ele_rec);
    case 1: return (int32_t)sizeof(bmx_particles);
    case 2: return (int32_t)sizeof(bmx_stats);
    case 3: return (int32_t)sizeof(bmx_grid);
    case 4: return (int32_t)sizeof(bmx_config);
    case 5: return (int32_t)sizeof(bmx_schedule);
    case 6: return (int32_t)sizeof(bmx_rad_map);
    case 7: return (int32_t)sizeof(bmx_radiation);
  }
  return -1;
}

int32_t bmx_ctx_create(bmx_ctx** out, int32_t device, uint32_t flags) {
  (void)flags;
  if (!out) return -1;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return -2;
  if (device < 0 || device >= n) return -3;
  if (hipSetDevice(device) != hipSuccess) return -4;
  bmx_ctx* c = new bmx_ctx();
  c->device = device;
  c->sched.turns_per_block = 32;
  c->sched.graph_depth = 1;
  c->sched.probe_pet_per_s = 0.0;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
    delete c;
    return -5;
  }
  c->scratch = scratch_address(device, &c->scratch_err);  // failure only disables bmx_track_bunch_noalloc
  *out = c;
  return 0;
}

void bmx_ctx_destroy(bmx_ctx* ctx) {
  if (!ctx) return;
  bmx_mlp_release(ctx);
  hipSetDevice(ctx->device);
  hipStreamSynchronize(ctx->stream);
  hipEventDestroy(ctx->ev0);
  hipEventDestroy(ctx->ev1);
  hipStreamDestroy(ctx->stream);
  delete ctx;
}

const char* bmx_last_error(const bmx_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

}  // extern "C"

// The element list of variant v: *p, *n, built in buf when it has to be
// (a worker thread's own buffer).
using VariantSource = std::function<void(int v, std::vector<bmx_ele_rec>& buf, const bmx_ele_rec** p, int* n)>;

// Compile, upload and set up every variant of a lattice (bmx_lattice_upload,
// bmx_lattice_upload_scaled).
static int32_t upload_impl(bmx_ctx* ctx, int32_t n_variants, const VariantSource& src, const bmx_ref* ref,
                           const bmx_config* cfg, bmx_lattice** out,
                           const std::function<const bmx_rad_map*(int v)>& rad_maps = nullptr,
                           const bmx_radiation* rad = nullptr) {
  bmx_config c = cfg ? *cfg : bmx_config{1, 0, 1e3, 0, 0};
  c.direction = 1;  // the forward program; the backward one of variant 0 is compiled below
  bmx_lattice* lat = new bmx_lattice();
  lat->device = ctx->device;
  lat->n_variants = n_variants;
  lat->abs_time = c.absolute_time_tracking != 0;
  // the variants compile independently, on up to 16 host threads, then
  // their programs are concatenated in variant order
  struct VarOut {
    std::vector<DevOp> prog;
    std::vector<int32_t> eb;
    std::vector<double> rad_table;
    LatticeConsts lc;
    std::string err;
    int n_ele = 0;
    double rf_freq = 0;
    bool ok = false;
  };
  const bool with_rad = rad && rad_maps && (rad->damping_on || rad->fluctuations_on);
  std::vector<VarOut> outs(n_variants);
  {
    std::atomic<int> next{0};
    auto worker = [&]() {
      std::vector<bmx_ele_rec> buf;
      for (int v; (v = next.fetch_add(1)) < n_variants;) {
        VarOut& o = outs[v];
        std::memset(&o.lc, 0, sizeof(o.lc));
        const bmx_ele_rec* e;
        int n;
        src(v, buf, &e, &n);
        o.n_ele = n;
        for (int i = 0; i < n; i++)
          if (e[i].key == BMX_KEY_RFCAVITY && e[i].is_on && e[i].rf_frequency != 0)
            o.rf_freq = o.rf_freq == 0 ? std::fabs(e[i].rf_frequency) : std::min(o.rf_freq, std::fabs(e[i].rf_frequency));
        RadInput ri;
        if (with_rad) {
          ri.maps = rad_maps(v);
          ri.cfg = *rad;
        }
        o.ok = compile_variant(e, n, *ref, c, o.prog, o.eb, &o.lc, &o.err, with_rad ? &ri : nullptr, &o.rad_table);
      }
    };
    unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    int nt = (int)std::min<unsigned>(std::min<unsigned>(hw, 16u), (unsigned)n_variants);
    std::vector<std::thread> th;
    for (int t = 1; t < nt; t++) th.emplace_back(worker);
    worker();
    for (auto& t : th) t.join();
  }
  std::vector<DevOp> prog;
  std::vector<double> rad_table;
  size_t total_ops = 0;
  for (const VarOut& o : outs) total_ops += o.prog.size();
  prog.reserve(total_ops);
  for (int v = 0; v < n_variants; v++) {
    VarOut& o = outs[v];
    if (!o.ok) {
      delete lat;
      return set_err(ctx, "variant " + std::to_string(v) + ": " + o.err);
    }
    int32_t base = (int32_t)prog.size();
    lat->var_op_begin.push_back(base);
    // each variant numbered its radiation maps from 0: shift them past the
    // maps of the variants before it
    const double rad_base = (double)(rad_table.size() / RAD_STRIDE);
    for (DevOp& op : o.prog)
      if (op.type == OP_RAD) op.c[0] += rad_base;
    rad_table.insert(rad_table.end(), o.rad_table.begin(), o.rad_table.end());
    lat->rad_fluct = lat->rad_fluct || o.lc.rad_fluct;
    prog.insert(prog.end(), o.prog.begin(), o.prog.end());
    for (int32_t& x : o.eb) x += base;
    LatticeConsts& lc = o.lc;
    lc.max_aperture_limit = c.max_aperture_limit;
    lc.two_max_ap = 2.0 * c.max_aperture_limit;
    lc.aperture_limit_on = c.aperture_limit_on;
    if (v == 0) lat->lc = lc;
    else if (lc.p0c != lat->lc.p0c || lc.mc2 != lat->lc.mc2) {
      delete lat;
      return set_err(ctx, "all variants must share the reference energy and species");
    }
    lat->ele_op_begin.push_back(std::move(o.eb));
    lat->var_n_ele.push_back(o.n_ele);
    lat->var_rf_freq.push_back(o.rf_freq);
    std::vector<DevOp>().swap(o.prog);
  }
  lat->var_op_begin.push_back((int32_t)prog.size());
  lat->n_ops = (int64_t)prog.size();
  CellPlan plan;
  bool cells = plan_cells(prog, &plan);
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  if (hipMalloc(&lat->d_ops, prog.size() * sizeof(DevOp)) != hipSuccess ||
      hipMalloc(&lat->d_var_op_begin, lat->var_op_begin.size() * sizeof(int32_t)) != hipSuccess ||
      hipMalloc(&lat->d_var_n_ele, lat->var_n_ele.size() * sizeof(int32_t)) != hipSuccess) {
    bmx_lattice_destroy(lat);
    return set_err(ctx, "hipMalloc of the element program failed");
  }
  HIP_TRY(ctx, hipMemcpy(lat->d_ops, prog.data(), prog.size() * sizeof(DevOp), hipMemcpyHostToDevice));
  HIP_TRY(ctx, hipMemcpy(lat->d_var_op_begin, lat->var_op_begin.data(), lat->var_op_begin.size() * sizeof(int32_t),
                         hipMemcpyHostToDevice));
  HIP_TRY(ctx, hipMemcpy(lat->d_var_n_ele, lat->var_n_ele.data(), lat->var_n_ele.size() * sizeof(int32_t),
                         hipMemcpyHostToDevice));
  if (!rad_table.empty()) {
    if (hipMalloc(&lat->d_rad, rad_table.size() * sizeof(double)) != hipSuccess) {
      bmx_lattice_destroy(lat);
      return set_err(ctx, "hipMalloc of the radiation maps failed");
    }
    HIP_TRY(ctx, hipMemcpy(lat->d_rad, rad_table.data(), rad_table.size() * sizeof(double), hipMemcpyHostToDevice));
    lat->lc.rad = lat->d_rad;
    lat->lc.rad_fluct = lat->rad_fluct ? 1 : 0;
    lat->lc.rad_key0 = (uint32_t)(rad->seed & 0xffffffffu);
    lat->lc.rad_key1 = (uint32_t)(rad->seed >> 32);
  }
  if (cells) {
    std::string jerr;
    const JitKernels* jk = jit_kernels(ctx->device, prog, plan, &jerr);
    if (jk && hipMalloc(&lat->d_cell_prv, plan.cell_prv.size() * sizeof(int32_t)) == hipSuccess &&
        hipMemcpy(lat->d_cell_prv, plan.cell_prv.data(), plan.cell_prv.size() * sizeof(int32_t),
                  hipMemcpyHostToDevice) == hipSuccess) {
      lat->jit = jk;
      lat->cell_len = plan.len;
      lat->n_cells = (int64_t)plan.n_cells;
      lat->lc.cell_prv = lat->d_cell_prv;
    } else if (!jk) {
      static std::once_flag once;
      std::call_once(once, [&] { std::fprintf(stderr, "bmx: cell kernels unavailable, using the interpreter: %s\n", jerr.c_str()); });
    }
  }
  std::vector<bmx_ele_rec> buf0;
  const bmx_ele_rec* e0;
  int n0;
  src(0, buf0, &e0, &n0);
  {  // reference times of variant 0's element boundaries (particle time t, loss edges)
    lat->rt.resize(n0 + 1);
    lat->rt[0] = e0[0].ref_time - e0[0].delta_ref_time;
    for (int j = 1; j <= n0; j++) lat->rt[j] = e0[j - 1].ref_time;
    if (hipMalloc(&lat->d_rt, lat->rt.size() * sizeof(double)) != hipSuccess ||
        hipMemcpy(lat->d_rt, lat->rt.data(), lat->rt.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
      bmx_lattice_destroy(lat);
      return set_err(ctx, "upload of the reference times failed");
    }
    lat->lc.rt = lat->d_rt;
  }
  {  // backward program of variant 0: elements last to first, each with direction -1
    bmx_config cb = c;
    cb.direction = -1;
    std::vector<DevOp> bprog;
    LatticeConsts lb;
    std::memset(&lb, 0, sizeof(lb));
    std::string berr;
    RadInput ri0;
    std::vector<double> brad;
    if (with_rad) {
      ri0.maps = rad_maps(0);
      ri0.cfg = *rad;
    }
    if (compile_variant(e0, n0, *ref, cb, bprog, lat->eb_bwd, &lb, &berr, with_rad ? &ri0 : nullptr, &brad)) {
      lb.max_aperture_limit = cb.max_aperture_limit;
      lb.two_max_ap = 2.0 * cb.max_aperture_limit;
      lb.aperture_limit_on = cb.aperture_limit_on;
      if (hipMalloc(&lat->d_ops_bwd, bprog.size() * sizeof(DevOp)) != hipSuccess) {
        lat->d_ops_bwd = nullptr;
        bmx_lattice_destroy(lat);
        return set_err(ctx, "hipMalloc of the backward program failed");
      }
      if (hipMemcpy(lat->d_ops_bwd, bprog.data(), bprog.size() * sizeof(DevOp), hipMemcpyHostToDevice) != hipSuccess) {
        bmx_lattice_destroy(lat);
        return set_err(ctx, "upload of the backward program failed");
      }
      lb.rt = lat->d_rt;
      lat->lc_bwd = lb;
    } else {
      lat->eb_bwd.clear();
      lat->bwd_err = berr;
    }
  }
  *out = lat;
  return 0;
}

extern "C" {

int32_t bmx_lattice_upload(bmx_ctx* ctx, const bmx_ele_rec* eles, int32_t n_ele_total, const int32_t* variant_offsets,
                           int32_t n_variants, const bmx_ref* ref, const bmx_config* cfg, bmx_lattice** out) {
  if (!ctx || !eles || !ref || !out || n_ele_total <= 0 || n_variants <= 0) return set_err(ctx, "invalid arguments");
  *out = nullptr;
  std::vector<int32_t> offs;
  if (variant_offsets) offs.assign(variant_offsets, variant_offsets + n_variants + 1);
  else if (n_variants == 1) offs = {0, n_ele_total};
  else return set_err(ctx, "variant_offsets is required for n_variants > 1");
  for (int v = 0; v < n_variants; v++)
    if (offs[v] < 0 || offs[v + 1] <= offs[v] || offs[v + 1] > n_ele_total) return set_err(ctx, "bad variant_offsets");
  return upload_impl(ctx, n_variants, [&](int v, std::vector<bmx_ele_rec>&, const bmx_ele_rec** p, int* n) {
    *p = eles + offs[v];
    *n = offs[v + 1] - offs[v];
  }, ref, cfg, out);
}

int32_t bmx_lattice_upload_rad(bmx_ctx* ctx, const bmx_ele_rec* eles, int32_t n_ele_total,
                               const int32_t* variant_offsets, int32_t n_variants, const bmx_rad_map* maps,
                               const bmx_radiation* rad, const bmx_ref* ref, const bmx_config* cfg,
                               bmx_lattice** out) {
  if (!ctx || !eles || !ref || !out || n_ele_total <= 0 || n_variants <= 0) return set_err(ctx, "invalid arguments");
  *out = nullptr;
  if (rad && (rad->damping_on || rad->fluctuations_on) && !maps)
    return set_err(ctx, "bmx_lattice_upload_rad: radiation switched on without maps");
  std::vector<int32_t> offs;
  if (variant_offsets) offs.assign(variant_offsets, variant_offsets + n_variants + 1);
  else if (n_variants == 1) offs = {0, n_ele_total};
  else return set_err(ctx, "variant_offsets is required for n_variants > 1");
  for (int v = 0; v < n_variants; v++)
    if (offs[v] < 0 || offs[v + 1] <= offs[v] || offs[v + 1] > n_ele_total) return set_err(ctx, "bad variant_offsets");
  return upload_impl(
      ctx, n_variants,
      [&](int v, std::vector<bmx_ele_rec>&, const bmx_ele_rec** p, int* n) {
        *p = eles + offs[v];
        *n = offs[v + 1] - offs[v];
      },
      ref, cfg, out, [&](int v) { return maps + 2 * (size_t)offs[v]; }, rad);
}

int32_t bmx_lattice_set_rad_stream(bmx_lattice* lat, uint64_t seed, uint64_t next_call) {
  if (!lat) return -1;
  lat->lc.rad_key0 = (uint32_t)(seed & 0xffffffffu);
  lat->lc.rad_key1 = (uint32_t)(seed >> 32);
  lat->rad_next_call.store(next_call);
  return 0;
}

int32_t bmx_lattice_upload_scaled(bmx_ctx* ctx, const bmx_ele_rec* eles, int32_t n_ele, int32_t n_variants,
                                  int32_t attr, const int32_t* sel, int32_t n_sel, const double* factors,
                                  const bmx_ref* ref, const bmx_config* cfg, bmx_lattice** out) {
  if (!ctx || !eles || !ref || !out || n_ele <= 0 || n_variants <= 0 || n_sel < 0 || (n_sel > 0 && (!sel || !factors)))
    return set_err(ctx, "invalid arguments");
  *out = nullptr;
  double bmx_ele_rec::*field = attr == BMX_ATTR_K1 ? &bmx_ele_rec::k1
                             : attr == BMX_ATTR_K2 ? &bmx_ele_rec::k2
                             : attr == BMX_ATTR_K3 ? &bmx_ele_rec::k3 : nullptr;
  if (!field) return set_err(ctx, "attr must be BMX_ATTR_K1, BMX_ATTR_K2 or BMX_ATTR_K3");
  for (int j = 0; j < n_sel; j++)
    if (sel[j] < 0 || sel[j] >= n_ele) return set_err(ctx, "selected element index out of range");
  return upload_impl(ctx, n_variants, [&](int v, std::vector<bmx_ele_rec>& buf, const bmx_ele_rec** p, int* n) {
    buf.assign(eles, eles + n_ele);
    const double* f = factors + (size_t)v * n_sel;
    for (int j = 0; j < n_sel; j++) buf[sel[j]].*field = buf[sel[j]].*field * f[j];
    *p = buf.data();
    *n = n_ele;
  }, ref, cfg, out);
}


int32_t bmx_lattice_check(const bmx_ele_rec* eles, int32_t n_ele, const bmx_ref* ref, const bmx_config* cfg,
                          int64_t* n_ops, char* msg, int32_t msg_len) {
  if (msg && msg_len > 0) msg[0] = '\0';
  std::string err;
  int32_t rc = 0;
  std::vector<DevOp> prog;
  if (!eles || !ref || n_ele <= 0) {
    err = "invalid arguments";
    rc = -1;
  } else {
    bmx_config c = cfg ? *cfg : bmx_config{1, 0, 1e3};
    std::vector<int32_t> eb;
    LatticeConsts lc;
    std::memset(&lc, 0, sizeof(lc));
    if (!compile_variant(eles, n_ele, *ref, c, prog, eb, &lc, &err)) rc = -1;
  }
  if (n_ops) *n_ops = rc == 0 ? (int64_t)prog.size() : 0;
  if (rc != 0 && msg && msg_len > 0) std::snprintf(msg, (size_t)msg_len, "%s", err.c_str());
  return rc;
}

int32_t bmx_lattice_cells(const bmx_lattice* lat, int32_t* cell_len, int64_t* n_cells) {
  if (!lat || !cell_len || !n_cells) return -1;
  *cell_len = lat->jit ? lat->cell_len : 0;
  *n_cells = lat->jit ? lat->n_cells : 0;
  return 0;
}

void bmx_lattice_destroy(bmx_lattice* lat) {
  if (!lat) return;
  hipSetDevice(lat->device);
  hipFree(lat->d_ops);
  hipFree(lat->d_var_op_begin);
  hipFree(lat->d_var_n_ele);
  if (lat->d_cell_prv) hipFree(lat->d_cell_prv);
  if (lat->d_ops_bwd) hipFree(lat->d_ops_bwd);
  if (lat->d_rt) hipFree(lat->d_rt);
  if (lat->d_rad) hipFree(lat->d_rad);
  delete lat;
}

}  // extern "C"

// Device views of a bmx_particles: host arrays are uploaded into buffers the
// call owns; results go back and the buffers are freed in finish_particles.
struct Staged {
  SoA s;
  void* dp[11];
  void* hp[11];
  size_t by[11];
  std::vector<void*> owned;
  unsigned long long* d_lost = nullptr;
  bool own_lost = true;  // d_lost was allocated by the call (not the static scratch)
};

static int stage_particles(bmx_ctx* ctx, const bmx_particles* p, Staged& st) {
  int64_t n = p->n;
  size_t bd = (size_t)n * sizeof(double), bi = (size_t)n * sizeof(int32_t);
  void* hp[11] = {p->x, p->px, p->y, p->py, p->z, p->pz, p->beta, p->state, p->ix_lost, p->turn_lost, p->t};
  size_t by[11] = {bd, bd, bd, bd, bd, bd, bd, bi, bi, bi, bd};
  int rc = 0;
  for (int k = 0; k < 11; k++) {
    st.hp[k] = hp[k];
    st.by[k] = by[k];
    st.dp[k] = nullptr;
    if (rc || !hp[k]) continue;
    if (p->on_device) { st.dp[k] = hp[k]; continue; }
    if (hipMalloc(&st.dp[k], by[k]) != hipSuccess) { rc = -1; continue; }
    st.owned.push_back(st.dp[k]);
    if (hipMemcpyAsync(st.dp[k], hp[k], by[k], hipMemcpyHostToDevice, ctx->stream) != hipSuccess) rc = -1;
  }
  if (rc == 0 && hipMalloc(&st.d_lost, sizeof(unsigned long long)) != hipSuccess) rc = -1;
  if (rc != 0) {
    for (void* q : st.owned) (void)hipFree(q);
    if (st.d_lost) (void)hipFree(st.d_lost);
    return set_err(ctx, "device allocation or upload of the particle arrays failed");
  }
  (void)hipMemsetAsync(st.d_lost, 0, sizeof(unsigned long long), ctx->stream);
  st.s.x = (double*)st.dp[0]; st.s.px = (double*)st.dp[1]; st.s.y = (double*)st.dp[2]; st.s.py = (double*)st.dp[3];
  st.s.z = (double*)st.dp[4]; st.s.pz = (double*)st.dp[5]; st.s.beta = (double*)st.dp[6];
  st.s.state = (int32_t*)st.dp[7]; st.s.ix_lost = (int32_t*)st.dp[8]; st.s.turn_lost = (int32_t*)st.dp[9];
  st.s.t = (double*)st.dp[10];
  return 0;
}

static int finish_particles(bmx_ctx* ctx, const bmx_particles* p, Staged& st, hipError_t le, bmx_stats* out) {
  unsigned long long h_lost = 0;
  (void)hipMemcpyAsync(&h_lost, st.d_lost, sizeof(h_lost), hipMemcpyDeviceToHost, ctx->stream);
  if (!p->on_device)
    for (int k = 0; k < 11; k++)
      if (st.hp[k]) (void)hipMemcpyAsync(st.hp[k], st.dp[k], st.by[k], hipMemcpyDeviceToHost, ctx->stream);
  hipError_t se = hipStreamSynchronize(ctx->stream);
  float ms = 0;
  (void)hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
  for (void* q : st.owned) (void)hipFree(q);
  if (st.own_lost) (void)hipFree(st.d_lost);
  if (le != hipSuccess) return set_err(ctx, std::string("kernel launch: ") + hipGetErrorString(le));
  if (se != hipSuccess) return set_err(ctx, std::string("stream sync: ") + hipGetErrorString(se));
  if (out) {
    out->n_lost = (int64_t)h_lost;
    out->kernel_ms = ms;
  }
  return 0;
}

// Device views of a bmx_particles in the static scratch: device arrays are
// used in place, host arrays (at most NOALLOC_MAX_PARTICLES) are copied into
// their scratch slots; the counter is cleared by a kernel.  No allocation.
static int stage_particles_static(bmx_ctx* ctx, const bmx_particles* p, Staged& st) {
  int64_t n = p->n;
  unsigned char* base = (unsigned char*)ctx->scratch;
  size_t bd = (size_t)n * sizeof(double), bi = (size_t)n * sizeof(int32_t);
  void* hp[11] = {p->x, p->px, p->y, p->py, p->z, p->pz, p->beta, p->state, p->ix_lost, p->turn_lost, p->t};
  size_t by[11] = {bd, bd, bd, bd, bd, bd, bd, bi, bi, bi, bd};
  const size_t N = (size_t)NOALLOC_MAX_PARTICLES;
  for (int k = 0; k < 11; k++) {
    st.hp[k] = hp[k];
    st.by[k] = by[k];
    st.dp[k] = nullptr;
    if (!hp[k]) continue;
    if (p->on_device) { st.dp[k] = hp[k]; continue; }
    size_t off = (k < 7) ? NOALLOC_HEAD + (size_t)k * N * sizeof(double)
                 : (k == 10) ? NOALLOC_HEAD + 7 * N * sizeof(double)  // t in the eighth double slot
                             : NOALLOC_HEAD + 8 * N * sizeof(double) + (size_t)(k - 7) * N * sizeof(int32_t);
    st.dp[k] = base + off;
    if (hipMemcpyAsync(st.dp[k], hp[k], by[k], hipMemcpyHostToDevice, ctx->stream) != hipSuccess)
      return set_err(ctx, "upload of the particle arrays into the static scratch failed");
  }
  st.d_lost = (unsigned long long*)base;
  st.own_lost = false;
  hipLaunchKernelGGL(bmx_zero_u64_kernel, dim3(1), dim3(1), 0, ctx->stream, st.d_lost);
  st.s.x = (double*)st.dp[0]; st.s.px = (double*)st.dp[1]; st.s.y = (double*)st.dp[2]; st.s.py = (double*)st.dp[3];
  st.s.z = (double*)st.dp[4]; st.s.pz = (double*)st.dp[5]; st.s.beta = (double*)st.dp[6];
  st.s.state = (int32_t*)st.dp[7]; st.s.ix_lost = (int32_t*)st.dp[8]; st.s.turn_lost = (int32_t*)st.dp[9];
  st.s.t = (double*)st.dp[10];
  return 0;
}

static bool particles_ok(const bmx_particles* p) {
  return p->x && p->px && p->y && p->py && p->z && p->pz && p->beta && p->state;
}

// bmx_track_bunch and bmx_track_bunch_noalloc: one body, the particle
// staging (allocated per call, or the static scratch) chosen by no_alloc.
static int32_t track_bunch_common(bmx_ctx* ctx, const bmx_lattice* lat, int32_t ix_start, int32_t ix_end,
                                  int32_t n_turns, int32_t direction, bmx_particles* p, bmx_stats* out,
                                  bool no_alloc) {
  if (!ctx || !lat || !p) return set_err(ctx, "invalid arguments");
  if (direction != 1 && direction != -1) return set_err(ctx, "direction must be +1 or -1");
  if (n_turns < 0) return set_err(ctx, "n_turns < 0");
  if (!particles_ok(p)) return set_err(ctx, "particle arrays x..pz, beta, state are required");
  int n_ele = lat->var_n_ele[0];
  if (ix_start < 0 || ix_start > n_ele || ix_end < 0 || ix_end > n_ele) return set_err(ctx, "ix_start/ix_end out of range");
  if (no_alloc) {
    if (!ctx->scratch) return set_err(ctx, "static scratch unavailable: " + ctx->scratch_err);
    if (!p->on_device && p->n > NOALLOC_MAX_PARTICLES)
      return set_err(ctx, "bmx_track_bunch_noalloc: " + std::to_string(p->n) +
                              " host-resident particles exceed the static staging area of " +
                              std::to_string(NOALLOC_MAX_PARTICLES) + " (pass device arrays, which have no limit)");
  }
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  if (direction == -1 && !lat->d_ops_bwd) return set_err(ctx, "backward tracking: " + lat->bwd_err);
  const bool bwd = direction == -1;
  const std::vector<int32_t>& eb = bwd ? lat->eb_bwd : lat->ele_op_begin[0];
  int lo1, hi1, lo2 = 0, hi2 = 0;
  if (!bwd) {
    // Bmad element numbering: elements ix_start+1 .. ix_end are array entries ix_start .. ix_end-1.
    if (ix_end > ix_start) { lo1 = eb[ix_start]; hi1 = eb[ix_end]; }
    else { lo1 = eb[ix_start]; hi1 = eb[n_ele]; lo2 = eb[0]; hi2 = eb[ix_end]; }
  } else {
    // backward (beam_mod.f90:131-147): elements ix_start down to ix_end+1,
    // wrapping through 1 and n_ele when ix_start <= ix_end; the backward
    // program holds array entry j at position n_ele - 1 - j
    if (ix_start > ix_end) { lo1 = eb[n_ele - ix_start]; hi1 = eb[n_ele - ix_end]; }
    else { lo1 = eb[n_ele - ix_start]; hi1 = eb[n_ele]; lo2 = eb[0]; hi2 = eb[n_ele - ix_end]; }
  }
  // reference time of the call (LatticeConsts::tt_*): the start edge, one
  // pass of the range (whole turns: t_lat) and the end edge of the last pass
  LatticeConsts lc = bwd ? lat->lc_bwd : lat->lc;
  {
    const std::vector<double>& rt = lat->rt;
    double T = lc.t_lat, range;
    if (!bwd) range = rt[ix_end] - rt[ix_start] + (ix_end <= ix_start ? T : 0.0);
    else range = rt[ix_start] - rt[ix_end] + (ix_start <= ix_end ? T : 0.0);
    if (ix_end == ix_start) range = T;
    lc.tt_sign = bwd ? -1 : 1;
    lc.tt_d = range - T;
    lc.ix_first = bwd ? ix_start - 1 : ix_start;
    lc.t_start = rt[ix_start];
    lc.t_end = rt[ix_start] + (double)lc.tt_sign * (double)n_turns * range;
  }
  rad_call_setup(lat, lc);
  int64_t n = p->n;
  if (out) std::memset(out, 0, sizeof(*out));
  if (n <= 0 || n_turns == 0) return 0;
  Staged st;
  if ((no_alloc ? stage_particles_static(ctx, p, st) : stage_particles(ctx, p, st)) != 0) return -1;
  const int block = 256;
  int64_t grid = (n + block * BMX_NP - 1) / (block * BMX_NP);
  (void)hipEventRecord(ctx->ev0, ctx->stream);
  if (bwd)  // the backward program runs on the op interpreter
    launch_track(nullptr, bmx_track_bunch_kernel, dim3((unsigned)grid), dim3(block), ctx->stream, lat->d_ops_bwd, lo1,
                 hi1, lo2, hi2, lc, n_turns, st.s, n, st.d_lost);
  else
    launch_track(lat->jit ? lat->jit->bunch : nullptr, bmx_track_bunch_kernel, dim3((unsigned)grid), dim3(block),
                 ctx->stream, lat->d_ops, lo1, hi1, lo2, hi2, lc, n_turns, st.s, n, st.d_lost);
  hipError_t le = hipGetLastError();
  (void)hipEventRecord(ctx->ev1, ctx->stream);
  return finish_particles(ctx, p, st, le, out);
}

extern "C" {

int32_t bmx_track_bunch(bmx_ctx* ctx, const bmx_lattice* lat, int32_t ix_start, int32_t ix_end, int32_t n_turns,
                        int32_t direction, bmx_particles* p, bmx_stats* out) {
  return track_bunch_common(ctx, lat, ix_start, ix_end, n_turns, direction, p, out, false);
}

int32_t bmx_track_bunch_noalloc(bmx_ctx* ctx, const bmx_lattice* lat, int32_t ix_start, int32_t ix_end,
                                int32_t n_turns, int32_t direction, bmx_particles* p, bmx_stats* out) {
  return track_bunch_common(ctx, lat, ix_start, ix_end, n_turns, direction, p, out, true);
}

int32_t bmx_track_multi(bmx_ctx* ctx, const bmx_lattice* lat, const int64_t* part_offsets, int32_t n_turns,
                        bmx_particles* p, bmx_stats* out) {
  if (!ctx || !lat || !p || !part_offsets) return set_err(ctx, "invalid arguments");
  if (n_turns < 0) return set_err(ctx, "n_turns < 0");
  if (!particles_ok(p)) return set_err(ctx, "particle arrays x..pz, beta, state are required");
  int nv = lat->n_variants;
  if (part_offsets[0] != 0 || part_offsets[nv] != p->n) return set_err(ctx, "part_offsets must run from 0 to p->n");
  int64_t max_count = 0;
  for (int v = 0; v < nv; v++) {
    if (part_offsets[v + 1] < part_offsets[v]) return set_err(ctx, "part_offsets must be non-decreasing");
    max_count = std::max(max_count, part_offsets[v + 1] - part_offsets[v]);
  }
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  if (out) std::memset(out, 0, sizeof(*out));
  if (p->n <= 0 || n_turns == 0) return 0;
  Staged st;
  bmx_particles q = *p;
  q.t = nullptr;  // the reference times are variant 0's: bmx_track_multi carries no particle time
  if (stage_particles(ctx, &q, st) != 0) return -1;
  int64_t* d_off = nullptr;
  if (hipMalloc(&d_off, (nv + 1) * sizeof(int64_t)) != hipSuccess) {
    finish_particles(ctx, p, st, hipSuccess, nullptr);
    return set_err(ctx, "hipMalloc of the variant offsets failed");
  }
  (void)hipMemcpyAsync(d_off, part_offsets, (nv + 1) * sizeof(int64_t), hipMemcpyHostToDevice, ctx->stream);
  const int block = 256;
  LatticeConsts lc = lat->lc;
  rad_call_setup(lat, lc);
  (void)hipEventRecord(ctx->ev0, ctx->stream);
  launch_track(lat->jit ? lat->jit->multi : nullptr, bmx_track_multi_kernel,
               dim3((unsigned)((max_count + block * BMX_NP - 1) / (block * BMX_NP)), (unsigned)nv), dim3(block),
               ctx->stream, lat->d_ops, lat->d_var_op_begin, d_off, lc, n_turns, st.s, st.d_lost);
  hipError_t le = hipGetLastError();
  (void)hipEventRecord(ctx->ev1, ctx->stream);
  int rc = finish_particles(ctx, p, st, le, out);
  (void)hipFree(d_off);
  return rc;
}

}  // extern "C"

// Survival tracking of per-variant particle stores in turn blocks with lane
// compaction (bmx_grid_block_kernel), shared by the polar grid and explicit
// initial conditions.  The caller fills store 0 and the survival array through
// `init` (a launcher run on the stream after the buffers exist).
struct SurvivalRun {
  int nv = 0, stride = 0;
  GridSoA st[2];
  int32_t* d_surv = nullptr;
  int32_t* d_count = nullptr;
  int32_t* d_turn0 = nullptr;
  int32_t* d_order = nullptr;  // variants in dispatch order
  unsigned long long* d_pet = nullptr;
  std::vector<void*> owned;
  ~SurvivalRun() {
    for (void* q : owned) (void)hipFree(q);
  }
  bool alloc(int n_var, int n_stride) {
    nv = n_var;
    stride = n_stride;
    int64_t total = (int64_t)nv * stride;
    auto dalloc = [&](size_t bytes) -> void* {
      void* p = nullptr;
      if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
      owned.push_back(p);
      return p;
    };
    for (int b = 0; b < 2; b++) {
      double** f[7] = {&st[b].x, &st[b].px, &st[b].y, &st[b].py, &st[b].z, &st[b].pz, &st[b].beta};
      for (int k = 0; k < 7; k++)
        if (!(*f[k] = (double*)dalloc(total * sizeof(double)))) return false;
      if (!(st[b].ic = (int32_t*)dalloc(total * sizeof(int32_t)))) return false;
    }
    return (d_surv = (int32_t*)dalloc(total * sizeof(int32_t))) && (d_count = (int32_t*)dalloc(2 * nv * sizeof(int32_t))) &&
           (d_turn0 = (int32_t*)dalloc(sizeof(int32_t))) && (d_order = (int32_t*)dalloc(nv * sizeof(int32_t))) &&
           (d_pet = (unsigned long long*)dalloc(sizeof(unsigned long long)));
  }
};

// Runs the turn blocks; counts[v] = initial particles of variant v.  Fills the
// host survival array (nv * stride), PET and timing.  Blocks alternate the
// parity p: store p -> store 1-p, counts p -> counts 1-p, then the advance
// kernel clears counts p and moves the first turn on.  graph_depth > 1
// captures that many (an even number of) steps into one hipGraph, replayed
// until the scan ends; the host reads the live counts after every replay and
// re-captures with a smaller launch once the live particles fall below 3/4 of
// the captured size.  graph_depth <= 1 launches and synchronises block by block.
// The variants are dispatched longest program first (blockIdx.y -> variant
// through d_order): with survivors of uneven variants left late in a scan,
// about 1.3 workgroups per slot, the long workgroups then run from the start
// of a block instead of after a round of short ones (6.6 -> 4.7 s on the
// 256-variant error study).
static int run_survival(bmx_ctx* ctx, const bmx_lattice* lat, SurvivalRun& R, std::vector<int32_t> h_count,
                        int n_turns, int tpb, int graph_depth, int32_t* h_surv, bmx_stats* out,
                        const std::function<void()>& init) {
  const int block = GRID_BLOCK;
  int nv = R.nv;
  if (graph_depth > 1 && (graph_depth & 1)) graph_depth++;
  (void)hipMemcpyAsync(R.d_count, h_count.data(), nv * sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream);
  (void)hipMemsetAsync(R.d_count + nv, 0, nv * sizeof(int32_t), ctx->stream);
  (void)hipMemsetAsync(R.d_pet, 0, sizeof(unsigned long long), ctx->stream);
  (void)hipMemsetAsync(R.d_turn0, 0, sizeof(int32_t), ctx->stream);
  // dispatch order: longest op program first (ties by index), so that a
  // block's long workgroups start first and the short ones fill in around them
  std::vector<int32_t> order(nv);
  for (int v = 0; v < nv; v++) order[v] = v;
  std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) {
    return lat->var_op_begin[a + 1] - lat->var_op_begin[a] > lat->var_op_begin[b + 1] - lat->var_op_begin[b];
  });
  (void)hipMemcpyAsync(R.d_order, order.data(), nv * sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream);
  (void)hipEventRecord(ctx->ev0, ctx->stream);
  init();
  int cur = 0, turn0 = 0, max_live = 0;
  for (int v = 0; v < nv; v++) max_live = std::max(max_live, h_count[v]);
  auto step = [&](int p, int live) {
    dim3 gd((unsigned)((live + block * BMX_NP - 1) / (block * BMX_NP)), (unsigned)nv);
    launch_track(lat->jit ? lat->jit->grid : nullptr, bmx_grid_block_kernel, gd, dim3(block), ctx->stream,
                 lat->d_ops, lat->d_var_op_begin, lat->d_var_n_ele, lat->lc, R.stride, R.st[p], R.st[1 - p],
                 R.d_count + p * nv, R.d_count + (1 - p) * nv, R.d_turn0, tpb, n_turns, R.d_surv, R.d_pet,
                 R.d_order);
    hipLaunchKernelGGL(bmx_grid_advance_kernel, dim3(1), dim3(256), 0, ctx->stream, R.d_count + p * nv, nv, R.d_turn0,
                       tpb);
  };
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int graph_live = 0;
  hipError_t le = hipGetLastError();
  while (le == hipSuccess && turn0 < n_turns && max_live > 0) {
    if (graph_depth > 1) {
      if (!exec || max_live < (graph_live * 3) / 4) {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        le = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal);
        if (le != hipSuccess) break;
        for (int k = 0; k < graph_depth; k++) step((cur + k) & 1, max_live);
        le = hipStreamEndCapture(ctx->stream, &graph);
        if (le == hipSuccess) le = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        if (le != hipSuccess) break;
        graph_live = max_live;
      }
      le = hipGraphLaunch(exec, ctx->stream);
      if (le != hipSuccess) break;
      // an even number of steps: the next block reads counts `cur` again
    } else {
      step(cur, max_live);
      le = hipGetLastError();
      if (le != hipSuccess) break;
      cur = 1 - cur;
    }
    (void)hipMemcpyAsync(h_count.data(), R.d_count + cur * nv, nv * sizeof(int32_t), hipMemcpyDeviceToHost,
                         ctx->stream);
    (void)hipMemcpyAsync(&turn0, R.d_turn0, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream);
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) break;
    max_live = 0;
    for (int v = 0; v < nv; v++) max_live = std::max(max_live, h_count[v]);
  }
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  (void)hipEventRecord(ctx->ev1, ctx->stream);
  unsigned long long h_pet = 0;
  int64_t total = (int64_t)nv * R.stride;
  (void)hipMemcpyAsync(&h_pet, R.d_pet, sizeof(h_pet), hipMemcpyDeviceToHost, ctx->stream);
  (void)hipMemcpyAsync(h_surv, R.d_surv, total * sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream);
  hipError_t se = hipStreamSynchronize(ctx->stream);
  float ms = 0;
  (void)hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
  if (le != hipSuccess) return set_err(ctx, std::string("turn blocks: ") + hipGetErrorString(le));
  if (se != hipSuccess) return set_err(ctx, std::string("stream sync: ") + hipGetErrorString(se));
  if (out) {
    std::memset(out, 0, sizeof(*out));
    out->kernel_ms = ms;
    out->pet = (double)h_pet;
  }
  return 0;
}

extern "C" {

int32_t bmx_track_grid(bmx_ctx* ctx, const bmx_lattice* lat, const bmx_grid* grid, const double* ref_orbit,
                       int32_t n_turns, int32_t* survival_turn, double* da_area, bmx_stats* out) {
  if (lat && lat->rad_fluct) return set_err(ctx, "bmx_track_grid: radiation fluctuations are tracked by bmx_track_bunch and bmx_track_multi only");
  if (!ctx || !lat || !grid || !survival_turn) return set_err(ctx, "invalid arguments");
  if (grid->n_angle <= 0 || grid->n_amp <= 0 || grid->n_pz <= 0 || n_turns < 0) return set_err(ctx, "bad grid");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  GridDev g;
  g.n_angle = grid->n_angle; g.n_amp = grid->n_amp; g.n_pz = grid->n_pz;
  g.n_ic = grid->n_angle * grid->n_amp * grid->n_pz;
  g.amp_max_x = grid->amp_max_x; g.amp_ratio_y = grid->amp_ratio_y;
  g.pz_min = grid->pz_min; g.pz_max = grid->pz_max;
  g.angle_min = grid->angle_min; g.angle_max = grid->angle_max;
  int nv = lat->n_variants;
  int tpb = grid->turns_per_block > 0 ? grid->turns_per_block : ctx->sched.turns_per_block;
  int64_t total = (int64_t)g.n_ic * nv;
  SurvivalRun R;
  double* d_ref = nullptr;
  if (!R.alloc(nv, g.n_ic)) return set_err(ctx, "device allocation for the grid failed");
  if (ref_orbit) {
    size_t bytes = (size_t)nv * g.n_pz * 6 * sizeof(double);
    if (hipMalloc(&d_ref, bytes) != hipSuccess) return set_err(ctx, "device allocation for the grid failed");
    R.owned.push_back(d_ref);
    (void)hipMemcpyAsync(d_ref, ref_orbit, bytes, hipMemcpyHostToDevice, ctx->stream);
  }
  std::vector<int32_t> h_count(nv, g.n_ic);
  const int block = 256;
  int rc = run_survival(ctx, lat, R, h_count, n_turns, tpb, ctx->sched.graph_depth, survival_turn, out, [&]() {
    hipLaunchKernelGGL(bmx_grid_init_kernel, dim3((unsigned)((g.n_ic + block - 1) / block), (unsigned)nv), dim3(block),
                       0, ctx->stream, g, lat->lc, R.st[0], d_ref, R.d_surv, n_turns);
  });
  if (rc != 0) return rc;
  // DA area per variant: on the middle pz slice, the boundary point of each
  // angle is the (x, y) offset of the last IC of the contiguous surviving run
  // from the smallest amplitude; area = the polygon through the origin and
  // those points in angle order, 1/2 sum (x_i y_{i+1} - x_{i+1} y_i).
  if (da_area) {
    int ipz = g.n_pz / 2;
    double dth = (g.n_angle > 1) ? (g.angle_max - g.angle_min) / (g.n_angle - 1) : 0.0;
    for (int v = 0; v < nv; v++) {
      std::vector<double> bx(g.n_angle, 0.0), by(g.n_angle, 0.0);
      for (int ia = 0; ia < g.n_angle; ia++) {
        int last = -1;
        for (int k = 0; k < g.n_amp; k++) {
          int ic = (ipz * g.n_angle + ia) * g.n_amp + k;
          if (survival_turn[(int64_t)v * g.n_ic + ic] >= n_turns) last = k;
          else break;
        }
        double amp = (last < 0) ? 0.0 : g.amp_max_x * (double)(last + 1) / g.n_amp;
        double th = g.angle_min + dth * ia;
        bx[ia] = amp * cos(th);
        by[ia] = amp * g.amp_ratio_y * sin(th);
      }
      double a = 0;
      for (int ia = 0; ia + 1 < g.n_angle; ia++) a += 0.5 * (bx[ia] * by[ia + 1] - bx[ia + 1] * by[ia]);
      da_area[v] = a;
    }
  }
  if (out) {
    int64_t live = 0;
    for (int64_t k = 0; k < total; k++) live += (survival_turn[k] >= n_turns);
    out->n_live = live;
    out->n_lost = total - live;
  }
  return 0;
}

int32_t bmx_track_ics(bmx_ctx* ctx, const bmx_lattice* lat, const double* ics, const int64_t* ic_offsets,
                      int32_t n_turns, int32_t turns_per_block, int32_t* survival_turn, bmx_stats* out) {
  if (lat && lat->rad_fluct) return set_err(ctx, "bmx_track_ics: radiation fluctuations are tracked by bmx_track_bunch and bmx_track_multi only");
  if (!ctx || !lat || !ics || !ic_offsets || !survival_turn || n_turns < 0) return set_err(ctx, "invalid arguments");
  int nv = lat->n_variants;
  if (ic_offsets[0] != 0) return set_err(ctx, "ic_offsets[0] must be 0");
  int64_t n = ic_offsets[nv], stride = 0;
  for (int v = 0; v < nv; v++) {
    if (ic_offsets[v + 1] < ic_offsets[v]) return set_err(ctx, "ic_offsets must be non-decreasing");
    stride = std::max(stride, ic_offsets[v + 1] - ic_offsets[v]);
  }
  if (stride > INT32_MAX / 2) return set_err(ctx, "too many initial conditions per variant");
  if (out) std::memset(out, 0, sizeof(*out));
  if (n == 0) return 0;
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  SurvivalRun R;
  if (!R.alloc(nv, (int)stride)) return set_err(ctx, "device allocation for the initial conditions failed");
  // padded host staging of x..pz, [v * stride + i]
  int64_t total = (int64_t)nv * stride;
  std::vector<double> pad(total);
  double* dst[6] = {R.st[0].x, R.st[0].px, R.st[0].y, R.st[0].py, R.st[0].z, R.st[0].pz};
  for (int c = 0; c < 6; c++) {
    std::fill(pad.begin(), pad.end(), 0.0);
    for (int v = 0; v < nv; v++)
      std::memcpy(&pad[(size_t)v * stride], ics + (size_t)c * n + ic_offsets[v],
                  (size_t)(ic_offsets[v + 1] - ic_offsets[v]) * sizeof(double));
    if (hipMemcpy(dst[c], pad.data(), total * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
      return set_err(ctx, "upload of the initial conditions failed");
  }
  std::vector<int32_t> h_count(nv);
  for (int v = 0; v < nv; v++) h_count[v] = (int32_t)(ic_offsets[v + 1] - ic_offsets[v]);
  std::vector<int32_t> h_surv(total);
  const int block = 256;
  int rc = run_survival(ctx, lat, R, h_count, n_turns, turns_per_block > 0 ? turns_per_block : ctx->sched.turns_per_block,
                        ctx->sched.graph_depth, h_surv.data(), out,
                        [&]() {
                          hipLaunchKernelGGL(bmx_ics_init_kernel,
                                             dim3((unsigned)((stride + block - 1) / block), (unsigned)nv), dim3(block),
                                             0, ctx->stream, lat->lc, R.st[0], R.d_count, (int)stride, R.d_surv, n_turns);
                        });
  if (rc != 0) return rc;
  int64_t live = 0;
  for (int v = 0; v < nv; v++)
    for (int64_t i = ic_offsets[v]; i < ic_offsets[v + 1]; i++) {
      survival_turn[i] = h_surv[(size_t)v * stride + (i - ic_offsets[v])];
      live += survival_turn[i] >= n_turns;
    }
  if (out) {
    out->n_live = live;
    out->n_lost = n - live;
  }
  return 0;
}



int32_t bmx_ctx_set_schedule(bmx_ctx* ctx, const bmx_schedule* s) {
  if (!ctx || !s || s->turns_per_block <= 0 || s->graph_depth < 0) return set_err(ctx, "invalid schedule");
  ctx->sched = *s;
  return 0;
}

int32_t bmx_ctx_get_schedule(const bmx_ctx* ctx, bmx_schedule* s) {
  if (!ctx || !s) return -1;
  *s = ctx->sched;
  return 0;
}

int32_t bmx_tune_grid(bmx_ctx* ctx, const bmx_lattice* lat, const bmx_grid* grid, const double* ref_orbit,
                      int32_t n_turns, bmx_schedule* best) {
  if (lat && lat->rad_fluct) return set_err(ctx, "bmx_tune_grid: radiation fluctuations are tracked by bmx_track_bunch and bmx_track_multi only");
  if (!ctx || !lat || !grid || !best || n_turns <= 0 || grid->n_angle <= 0 || grid->n_amp <= 0 || grid->n_pz <= 0)
    return set_err(ctx, "invalid arguments");
  // sparse sample of the grid: every 4th angle and amplitude, every pz slice, every variant
  int nv = lat->n_variants;
  std::vector<std::array<double, 3>> per;  // (x offset, y offset, pz slice)
  double dth = (grid->n_angle > 1) ? (grid->angle_max - grid->angle_min) / (grid->n_angle - 1) : 0.0;
  for (int ip = 0; ip < grid->n_pz; ip++)
    for (int ia = 0; ia < grid->n_angle; ia += 4)
      for (int k = 0; k < grid->n_amp; k += 4) {
        double amp = grid->amp_max_x * (double)(k + 1) / grid->n_amp, th = grid->angle_min + dth * ia;
        per.push_back({amp * cos(th), amp * grid->amp_ratio_y * sin(th), (double)ip});
      }
  int64_t m = (int64_t)per.size(), n = m * nv;
  std::vector<double> ics(6 * n);
  std::vector<int64_t> off(nv + 1);
  for (int v = 0; v <= nv; v++) off[v] = (int64_t)v * m;
  for (int v = 0; v < nv; v++)
    for (int64_t j = 0; j < m; j++) {
      int ip = (int)per[j][2];
      double pz0 = (grid->n_pz > 1) ? grid->pz_min + (grid->pz_max - grid->pz_min) * ip / (grid->n_pz - 1) : grid->pz_min;
      double r[6] = {0, 0, 0, 0, 0, pz0};
      if (ref_orbit)
        for (int c = 0; c < 6; c++) r[c] = ref_orbit[((int64_t)v * grid->n_pz + ip) * 6 + c];
      r[0] += per[j][0];
      r[2] += per[j][1];
      for (int c = 0; c < 6; c++) ics[(size_t)c * n + off[v] + j] = r[c];
    }
  std::vector<int32_t> surv(n);
  const int tpbs[6] = {4, 8, 16, 32, 64, 128};
  const int depths[2] = {1, 8};
  int probe = std::min(n_turns, 64);
  bmx_schedule saved = ctx->sched, pick = saved;
  double best_rate = -1;
  for (int d : depths)
    for (int t : tpbs) {
      if (t > probe && t != tpbs[0]) continue;
      ctx->sched.turns_per_block = t;
      ctx->sched.graph_depth = d;
      bmx_stats st;
      auto t0 = std::chrono::steady_clock::now();
      int rc = bmx_track_ics(ctx, lat, ics.data(), off.data(), probe, t, surv.data(), &st);
      double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      if (rc != 0) {
        ctx->sched = saved;
        return rc;
      }
      double rate = st.pet / std::max(wall, 1e-9);
      if (rate > best_rate) {
        best_rate = rate;
        pick.turns_per_block = t;
        pick.graph_depth = d;
        pick.probe_pet_per_s = rate;
      }
    }
  ctx->sched = pick;
  *best = pick;
  return 0;
}

int32_t bmx_bunch_moments(bmx_ctx* ctx, const bmx_particles* p, double* moments) {
  if (!ctx || !p || !moments || !particles_ok(p)) return set_err(ctx, "invalid arguments");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  for (int k = 0; k < N_MOM; k++) moments[k] = 0.0;
  if (p->n <= 0) return 0;
  Staged st;
  bmx_particles q = *p;
  q.ix_lost = nullptr;
  q.turn_lost = nullptr;
  q.t = nullptr;
  if (stage_particles(ctx, &q, st) != 0) return -1;
  const int blocks = 1024;
  double *d_part = nullptr, *d_out = nullptr;
  if (hipMalloc(&d_part, (size_t)blocks * N_MOM * sizeof(double)) != hipSuccess ||
      hipMalloc(&d_out, N_MOM * sizeof(double)) != hipSuccess) {
    if (d_part) (void)hipFree(d_part);
    st.hp[0] = nullptr;  // nothing to copy back
    q.on_device = 1;
    finish_particles(ctx, &q, st, hipSuccess, nullptr);
    return set_err(ctx, "hipMalloc for the moment reduction failed");
  }
  size_t lds = (size_t)N_MOM * 256 * sizeof(double);
  hipLaunchKernelGGL(bmx_moments_partial_kernel, dim3(blocks), dim3(256), lds, ctx->stream, st.s, p->n, d_part);
  hipLaunchKernelGGL(bmx_moments_final_kernel, dim3(1), dim3(256), lds, ctx->stream, d_part, blocks, d_out);
  hipError_t le = hipGetLastError();
  (void)hipMemcpyAsync(moments, d_out, N_MOM * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
  // no coordinates change: skip the copy-back of host arrays
  q.on_device = 1;
  int rc = finish_particles(ctx, &q, st, le, nullptr);
  (void)hipFree(d_part);
  (void)hipFree(d_out);
  return rc;
}

// Spectral radius of a 6x6 matrix from the growth of its powers: M is
// squared K times with the largest entry renormalised to 1 each time, and
// log rho = log ||M^(2^K)|| / 2^K to within log(cond M) / 2^K.
static double spectral_radius6(const double (&M)[6][6]) {
  double A[6][6], T[6][6];
  double nrm = 0;
  for (int i = 0; i < 6; i++)
    for (int j = 0; j < 6; j++) nrm = std::max(nrm, std::fabs(M[i][j]));
  if (nrm == 0 || !std::isfinite(nrm)) return nrm;
  for (int i = 0; i < 6; i++)
    for (int j = 0; j < 6; j++) A[i][j] = M[i][j] / nrm;
  double lg = std::log(nrm);  // M^(2^k) = exp(lg) A
  const int K = 30;
  for (int k = 0; k < K; k++) {
    double n2 = 0;
    for (int i = 0; i < 6; i++)
      for (int j = 0; j < 6; j++) {
        double a = 0;
        for (int q = 0; q < 6; q++) a += A[i][q] * A[q][j];
        T[i][j] = a;
        n2 = std::max(n2, std::fabs(a));
      }
    if (n2 == 0) return 0;
    for (int i = 0; i < 6; i++)
      for (int j = 0; j < 6; j++) A[i][j] = T[i][j] / n2;
    lg = 2 * lg + std::log(n2);
  }
  return std::exp(lg / (double)(1 << K));
}

// closed_orbit_calc (bmad/code/closed_orbit_calc.f90:81-...) in the form the
// batched tracker serves: Newton iteration x <- x + (I - M)^-1 (T(x) - x) on
// the one-turn map T of every variant at once, with M from one-sided finite
// differences (one launch tracks 1 + n_dim probes per variant).  n_dim = 4
// keeps pz (and z) of the guess fixed; n_dim = 6 closes all six (rf on) and,
// as the reference does (:241-256, :400-410, :491-520): keeps z within half
// an rf wavelength of zero (the lowest rf frequency of the cavities that are
// on), limits a z step to a tenth of the wavelength, and, when the orbit
// found is longitudinally unstable (a one-turn eigenvalue beyond 1 + 1e-6),
// moves z by multiples of an eighth of the wavelength to the start whose
// one-turn matrix is the most stable and iterates again from there.  An
// orbit has converged when a step is below 1e-15 (relative to 1 + |x|) or,
// below 1e-9, no longer shrinks: the rounding floor of the one-turn map.
int32_t bmx_closed_orbit(bmx_ctx* ctx, const bmx_lattice* lat, int32_t n_dim, double* orbit, int32_t* status,
                         bmx_stats* out) {
  if (lat && lat->rad_fluct) return set_err(ctx, "bmx_closed_orbit: radiation fluctuations are tracked by bmx_track_bunch and bmx_track_multi only");
  if (!ctx || !lat || !orbit || !status || (n_dim != 4 && n_dim != 6)) return set_err(ctx, "invalid arguments");
  const int nv = lat->n_variants;
  const int np = 1 + n_dim;
  const double h = 1e-8;
  const double mc2 = lat->lc.mc2, p0c = lat->lc.p0c;
  std::vector<double> wavelen(nv, 0.0);
  if (n_dim == 6)
    for (int v = 0; v < nv; v++)
      if (lat->var_rf_freq[v] > 0) wavelen[v] = D_C_LIGHT * (p0c / std::sqrt(p0c * p0c + mc2 * mc2)) / lat->var_rf_freq[v];
  // x folded into [-amp, amp) (modulo2, sim_utils/special_functions/modulo2_mod.f90:85-99)
  auto modulo2 = [](double x, double amp) {
    double m = x - 2 * amp * std::floor(x / (2 * amp));
    return m >= amp ? m - 2 * amp : m;
  };
  if (out) std::memset(out, 0, sizeof(*out));

  // One tracking launch over sets of probes: each base orbit of a variant
  // with its n_dim displaced copies.  Per base: lost, T(base) and the
  // finite-difference Jacobian (columns past n_dim are the identity's).
  struct Probe {
    bool lost;
    double t[6];
    double m[6][6];
  };
  auto run_probes = [&](const std::vector<std::vector<std::array<double, 6>>>& bases,
                        std::vector<std::vector<Probe>>& res) -> int {
    std::vector<int64_t> off(nv + 1, 0);
    for (int v = 0; v < nv; v++) off[v + 1] = off[v] + (int64_t)bases[v].size() * np;
    const int64_t n = off[nv];
    res.assign(nv, {});
    if (n == 0) return 0;
    std::vector<double> c(6 * n), beta(n);
    std::vector<int32_t> state(n), ixl(n), tl(n);
    for (int v = 0; v < nv; v++)
      for (size_t b = 0; b < bases[v].size(); b++)
        for (int j = 0; j < np; j++) {
          int64_t k = off[v] + (int64_t)b * np + j;
          for (int d = 0; d < 6; d++) c[(size_t)d * n + k] = bases[v][b][d] + ((j > 0 && d == j - 1) ? h : 0.0);
          double pz = c[(size_t)5 * n + k], pc = p0c * (1.0 + pz);
          beta[k] = pc / std::sqrt(pc * pc + mc2 * mc2);
          state[k] = BMX_ALIVE;
        }
    bmx_particles p;
    std::memset(&p, 0, sizeof(p));
    p.x = &c[0]; p.px = &c[(size_t)n]; p.y = &c[(size_t)2 * n]; p.py = &c[(size_t)3 * n];
    p.z = &c[(size_t)4 * n]; p.pz = &c[(size_t)5 * n];
    p.beta = beta.data(); p.state = state.data(); p.ix_lost = ixl.data(); p.turn_lost = tl.data();
    p.n = n;
    p.on_device = 0;
    bmx_stats st;
    if (bmx_track_multi(ctx, lat, off.data(), 1, &p, &st) != 0) return -1;
    if (out) out->kernel_ms += st.kernel_ms;
    for (int v = 0; v < nv; v++)
      for (size_t b = 0; b < bases[v].size(); b++) {
        int64_t k0 = off[v] + (int64_t)b * np;
        Probe pr;
        pr.lost = false;
        for (int j = 0; j < np; j++) pr.lost |= state[k0 + j] != BMX_ALIVE;
        for (int d = 0; d < 6; d++) {
          pr.t[d] = c[(size_t)d * n + k0];
          for (int j = 0; j < 6; j++)
            pr.m[d][j] = j < n_dim ? (c[(size_t)d * n + k0 + 1 + j] - c[(size_t)d * n + k0]) / h : (d == j ? 1.0 : 0.0);
        }
        res[v].push_back(pr);
      }
    return 0;
  };

  std::vector<int> done(nv, 0), z_moves(nv, 0);
  std::vector<double> last_step(nv, 0.0), rho_found(nv, 0.0);
  for (int v = 0; v < nv; v++) {
    status[v] = 1;
    if (n_dim == 6 && wavelen[v] > 0) orbit[v * 6 + 4] = modulo2(orbit[v * 6 + 4], wavelen[v] / 2);
  }
  std::vector<std::vector<std::array<double, 6>>> bases(nv);
  std::vector<std::vector<Probe>> res;
  for (int it = 0; it < 60; it++) {
    for (int v = 0; v < nv; v++) {
      bases[v].clear();
      if (!done[v]) bases[v].push_back({orbit[v * 6], orbit[v * 6 + 1], orbit[v * 6 + 2], orbit[v * 6 + 3],
                                        orbit[v * 6 + 4], orbit[v * 6 + 5]});
    }
    if (run_probes(bases, res) != 0) return -1;
    // variants whose converged orbit is longitudinally unstable: the z search
    std::vector<int> search;
    for (int v = 0; v < nv; v++) {
      if (done[v]) continue;
      const Probe& pr = res[v][0];
      if (pr.lost) { status[v] = 2; done[v] = 1; continue; }
      // residual r = T(x) - x and (I - M) over the first n_dim coordinates
      double r[6], A[6][7];
      double rmax = 0;
      for (int d = 0; d < n_dim; d++) {
        r[d] = pr.t[d] - orbit[v * 6 + d];
        rmax = std::max(rmax, std::fabs(r[d]));
      }
      bool conv = rmax < 1e-15;
      if (!conv) {
        for (int d = 0; d < n_dim; d++) {
          for (int j = 0; j < n_dim; j++) A[d][j] = (d == j ? 1.0 : 0.0) - pr.m[d][j];
          A[d][n_dim] = r[d];
        }
        // Gaussian elimination with partial pivoting: (I - M) dx = r
        bool sing = false;
        for (int col = 0; col < n_dim && !sing; col++) {
          int piv = col;
          for (int row = col + 1; row < n_dim; row++)
            if (std::fabs(A[row][col]) > std::fabs(A[piv][col])) piv = row;
          if (std::fabs(A[piv][col]) < 1e-14) { sing = true; break; }
          for (int k = 0; k <= n_dim; k++) std::swap(A[col][k], A[piv][k]);
          for (int row = col + 1; row < n_dim; row++) {
            double f = A[row][col] / A[col][col];
            for (int k = col; k <= n_dim; k++) A[row][k] -= f * A[col][k];
          }
        }
        if (sing) { status[v] = 3; done[v] = 1; continue; }
        double dx[6];
        for (int row = n_dim - 1; row >= 0; row--) {
          double acc = A[row][n_dim];
          for (int k = row + 1; k < n_dim; k++) acc -= A[row][k] * dx[k];
          dx[row] = acc / A[row][row];
        }
        if (n_dim == 6 && wavelen[v] > 0) {  // a z step of at most a tenth of the wavelength
          double lim = std::fabs(dx[4]) / (wavelen[v] / 10);
          if (lim > 1)
            for (int d = 0; d < 6; d++) dx[d] /= lim;
        }
        double step = 0;
        for (int d = 0; d < n_dim; d++) {
          step = std::max(step, std::fabs(dx[d]) / (1.0 + std::fabs(orbit[v * 6 + d])));
          orbit[v * 6 + d] += dx[d];
        }
        if (n_dim == 6 && wavelen[v] > 0) orbit[v * 6 + 4] = modulo2(orbit[v * 6 + 4], wavelen[v] / 2);
        conv = step < 1e-15 || (it > 0 && step < 1e-9 && step > 0.5 * last_step[v]);
        last_step[v] = step;
      }
      if (!conv) continue;
      status[v] = 0;
      done[v] = 1;
      if (n_dim == 6 && wavelen[v] > 0 && z_moves[v] < 3) {
        rho_found[v] = spectral_radius6(pr.m);
        if (rho_found[v] > 1.000001) search.push_back(v);
      }
    }
    if (!search.empty()) {
      // candidates z0 + j lambda / 8, j = -3..4 (j = 0 is the orbit found)
      for (int v = 0; v < nv; v++) bases[v].clear();
      for (int v : search)
        for (int j = -3; j <= 4; j++) {
          if (j == 0) continue;
          std::array<double, 6> b = {orbit[v * 6], orbit[v * 6 + 1], orbit[v * 6 + 2],
                                     orbit[v * 6 + 3], orbit[v * 6 + 4] + j * wavelen[v] / 8, orbit[v * 6 + 5]};
          bases[v].push_back(b);
        }
      if (run_probes(bases, res) != 0) return -1;
      for (int v : search) {
        int jb = 0, q = 0;
        double rho0 = rho_found[v];
        for (int j = -3; j <= 4; j++) {
          if (j == 0) continue;
          const Probe& pr = res[v][q++];
          if (pr.lost) continue;
          double rho = spectral_radius6(pr.m);
          if (rho < rho0) { rho0 = rho; jb = j; }
        }
        if (jb != 0) {
          orbit[v * 6 + 4] = modulo2(orbit[v * 6 + 4] + jb * wavelen[v] / 8, wavelen[v] / 2);
          z_moves[v]++;
          status[v] = 1;
          done[v] = 0;
          last_step[v] = 0;
        }
      }
    }
    bool open = false;
    for (int v = 0; v < nv; v++) open |= !done[v];
    if (!open) break;
  }
  return 0;
}

int32_t bmx_one_turn_matrix(bmx_ctx* ctx, const bmx_lattice* lat, const double* orbit, const double* d_orb,
                            double* mat6, int32_t* status, bmx_stats* out) {
  if (lat && lat->rad_fluct) return set_err(ctx, "bmx_one_turn_matrix: radiation fluctuations are tracked by bmx_track_bunch and bmx_track_multi only");
  if (!ctx || !lat || !orbit || !mat6 || !status) return set_err(ctx, "invalid arguments");
  double d[6];
  for (int k = 0; k < 6; k++) {
    d[k] = d_orb ? d_orb[k] : 1e-5;  // bmad_com%d_orb
    if (!(d[k] > 0)) return set_err(ctx, "d_orb must be positive");
  }
  int nv = lat->n_variants;
  const int np = 12;  // orbit +- d[k] e_k
  int64_t n = (int64_t)nv * np;
  std::vector<int64_t> off(nv + 1);
  for (int v = 0; v <= nv; v++) off[v] = (int64_t)v * np;
  std::vector<double> c(6 * n), beta(n);
  std::vector<int32_t> state(n), ixl(n), tl(n);
  double mc2 = lat->lc.mc2, p0c = lat->lc.p0c;
  for (int v = 0; v < nv; v++)
    for (int j = 0; j < np; j++) {
      int64_t k = off[v] + j;
      for (int q = 0; q < 6; q++) {
        double x = orbit[v * 6 + q];
        if (q == j / 2) x += (j % 2 == 0) ? d[q] : -d[q];
        c[(size_t)q * n + k] = x;
      }
      double pz = c[(size_t)5 * n + k], pc = p0c * (1.0 + pz);
      beta[k] = pc / std::sqrt(pc * pc + mc2 * mc2);
      state[k] = BMX_ALIVE;
    }
  bmx_particles p;
  std::memset(&p, 0, sizeof(p));
  p.x = &c[0]; p.px = &c[(size_t)n]; p.y = &c[(size_t)2 * n]; p.py = &c[(size_t)3 * n];
  p.z = &c[(size_t)4 * n]; p.pz = &c[(size_t)5 * n];
  p.beta = beta.data(); p.state = state.data(); p.ix_lost = ixl.data(); p.turn_lost = tl.data();
  p.n = n;
  p.on_device = 0;
  bmx_stats st;
  if (bmx_track_multi(ctx, lat, off.data(), 1, &p, &st) != 0) return -1;
  if (out) *out = st;
  for (int v = 0; v < nv; v++) {
    bool lost = false;
    for (int j = 0; j < np; j++) lost |= state[off[v] + j] != BMX_ALIVE;
    status[v] = lost ? 2 : 0;
    for (int i = 0; i < 6; i++)
      for (int q = 0; q < 6; q++) {
        int64_t kp = off[v] + 2 * q, km = kp + 1;
        mat6[(size_t)v * 36 + i * 6 + q] =
            lost ? 0.0 : (c[(size_t)i * n + kp] - c[(size_t)i * n + km]) / (2.0 * d[q]);
      }
  }
  return 0;
}

int32_t bmx_transfer_matrix(bmx_ctx* ctx, const bmx_lattice* lat, int32_t ix_start, int32_t ix_end, int32_t n_turns,
                            int64_t n, const double* vec_in, double* vec_out, double* mat6, int32_t* state) {
  if (lat && lat->rad_fluct) return set_err(ctx, "bmx_transfer_matrix: radiation fluctuations are tracked by bmx_track_bunch and bmx_track_multi only");
  if (!ctx || !lat || !vec_in || !vec_out || !mat6 || !state || n < 0 || n_turns < 0)
    return set_err(ctx, "invalid arguments");
  if (lat->abs_time) return set_err(ctx, "bmx_transfer_matrix: lattices with absolute-time rf are not supported");
  int n_ele = lat->var_n_ele[0];
  if (ix_start < 0 || ix_start > n_ele || ix_end < 0 || ix_end > n_ele) return set_err(ctx, "ix_start/ix_end out of range");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  const std::vector<int32_t>& eb = lat->ele_op_begin[0];
  int lo1, hi1, lo2 = 0, hi2 = 0;
  if (ix_end > ix_start) { lo1 = eb[ix_start]; hi1 = eb[ix_end]; }
  else { lo1 = eb[ix_start]; hi1 = eb[n_ele]; lo2 = eb[0]; hi2 = eb[ix_end]; }
  if (n == 0 || n_turns == 0) {
    for (int64_t i = 0; i < n; i++) {
      state[i] = BMX_ALIVE;
      for (int r = 0; r < 6; r++) {
        vec_out[i * 6 + r] = vec_in[i * 6 + r];
        for (int q = 0; q < 6; q++) mat6[i * 36 + r * 6 + q] = r == q ? 1.0 : 0.0;
      }
    }
    return 0;
  }
  double* d = nullptr;
  size_t nb = (size_t)n * (6 + 6 + 36) * sizeof(double) + (size_t)n * sizeof(int32_t);
  HIP_TRY(ctx, hipMalloc(&d, nb));
  double *d_in = d, *d_out = d + 6 * n, *d_mat = d + 12 * n;
  int32_t* d_state = reinterpret_cast<int32_t*>(d + 48 * n);
  hipError_t e = hipMemcpyAsync(d_in, vec_in, (size_t)n * 6 * sizeof(double), hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess)
    e = launch_transfer(ctx->stream, lat->d_ops, lo1, hi1, lo2, hi2, lat->lc, n_turns, d_in, d_out, d_mat, d_state, n);
  if (e == hipSuccess)
    e = hipMemcpyAsync(vec_out, d_out, (size_t)n * 6 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(mat6, d_mat, (size_t)n * 36 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(state, d_state, (size_t)n * sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  (void)hipFree(d);
  if (e != hipSuccess) return set_err(ctx, std::string("bmx_transfer_matrix: ") + hipGetErrorString(e));
  return 0;
}

}  // extern "C"
