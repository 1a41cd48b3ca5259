// bmx_kernels.h -- device side of the tracker, part 2: the LDS-staged walk
// over the element program and the kernels.  Included after bmx_ops.h by
// bmx_track.hip and, as text, by the runtime-compiled cell kernels, which
// define BMX_CELL_LEN and jit_cell() in between.
#pragma once

// Element programs are staged through LDS: the workgroup copies OPS_CHUNK
// records at a time and every wave reads them at wave-uniform LDS
// addresses, instead of each CU fetching every record from L2 through its
// scalar cache on the dependent path of the element loop.  256 records are
// 32 KiB: four workgroups of 256 lanes (occupancy 4) fit in a CU's 160 KiB.
#ifndef BMX_OPS_CHUNK
#define BMX_OPS_CHUNK 256
#endif
constexpr int OPS_CHUNK = BMX_OPS_CHUNK;

// Particles per lane (BMX_NP, bmx_internal.h).  Each lane carries NP
// particles through the element program together, so the op dispatch, the
// LDS reads of the op constants and the loop control are paid once for NP
// particles, and the NP particle updates of an op sit in one basic block
// where they can be interleaved.
constexpr int GRID_BLOCK = BMX_GRID_BLOCK;

// The particles of one lane.  live[j]: still tracked.  A particle is lost at
// the op whose step set its state: on_loss() gets its coordinates at that
// moment, then it is parked at the reference orbit (coordinates zero, alive
// state) so that it keeps taking the ordinary paths of every op while the
// lane's other particles go on; its results are never read again.
template <int NP>
struct Pack {
  Lane L[NP];
  bool live[NP];
  int ie_lost[NP], lost_turn[NP];
  __device__ bool any() const {
    bool a = false;
    for (int j = 0; j < NP; j++) a |= live[j];
    return a;
  }
};

__device__ __forceinline__ void park(Lane& L, const LatticeConsts& lc) {
  L.x = 0.0; L.px = 0.0; L.y = 0.0; L.py = 0.0; L.z = 0.0; L.pz = 0.0;
  L.beta = lc.beta0;
  L.irp = 1.0;
  L.bb = 0.0;
#if BMX_ABS_TIME
  L.dt = 0.0;
#endif
#if BMX_RAD
  L.pid = 0;
#endif
  L.state = BMX_ALIVE;
}

// Reference time at the far edge, in the tracking direction, of array
// element ie on pass `turn` of a bmx_track_bunch call (variant 0).
__device__ __forceinline__ double edge_time(const LatticeConsts& lc, int ie, int turn) {
  bool fwd = lc.tt_sign > 0;
  double base = fwd ? lc.rt[ie + 1] : lc.rt[ie];
  bool wrapped = fwd ? (ie < lc.ix_first) : (ie > lc.ix_first);
  double k = (double)lc.tt_sign * ((double)turn + (wrapped ? 1.0 : 0.0));
  return base + k * lc.t_lat + (double)(lc.tt_sign * turn) * lc.tt_d;
}

// Walk of `turns` turns over ops [lo1, hi1) then [lo2, hi2) for one lane's
// particles.  The whole workgroup calls it together (it holds barriers); the
// element program comes through LDS one chunk at a time.  Lanes without a
// live particle sit out each chunk; a lane whose last particle is lost
// mid-chunk carries the parked particle to the end of the chunk (an exit
// test per op costs more than that).
//
// In a runtime-compiled cell kernel (BMX_CELL_LEN defined, with jit_cell()
// holding the cell's ops with their types and flags as constants), an op
// flagged F_CELL starts a run of BMX_CELL_LEN ops of the lattice's repeating
// cell, which jit_cell() tracks straight through; every other op goes
// through run_op.  A chunk ends before a cell that would cross its end
// (lc.cell_prv[i]: the last cell start at or before op i); a cell that
// crosses the end of the tracked range runs op by op.
//
// turn0: the pass number of the first turn (survival scans in turn blocks);
// the reference time of each pass (TurnTime, absolute-time rf) follows from
// it, the turn and whether the segment lies past the lattice end.
template <int NP, int CHUNK, class OnLoss>
__device__ __forceinline__ void walk(Pack<NP>& P, const DevOp* __restrict__ ops, int lo1, int hi1, int lo2, int hi2,
                                     int turns, int turn0, const LatticeConsts& lc, DevOp* sops, OnLoss on_loss) {
  for (int t = 0; t < turns; t++) {
    for (int seg = 0; seg < 2; seg++) {
      int lo = seg ? lo2 : lo1, hi = seg ? hi2 : hi1;
      TurnTime tt;
      tt.k = lc.tt_sign * (turn0 + t + seg);
      tt.r = (double)(lc.tt_sign * (turn0 + t)) * lc.tt_d;
      tt.turn = turn0 + t;
      for (int c = lo; c < hi;) {
        int m = min(CHUNK, hi - c);
#ifdef BMX_CELL_LEN
        {
          int s = lc.cell_prv[c + m - 1];
          if (s > c && s + BMX_CELL_LEN > c + m && s + BMX_CELL_LEN <= hi) m = s - c;
        }
#endif
        // barrier: the previous chunk is consumed; leave together once the workgroup has no live particle
        if (!__syncthreads_or(P.any())) return;
        const uint4* src = reinterpret_cast<const uint4*>(ops + c);
        uint4* dst = reinterpret_cast<uint4*>(sops);
        constexpr int W = (int)(sizeof(DevOp) / sizeof(uint4));
        for (int e = threadIdx.x; e < m * W; e += blockDim.x) dst[e] = src[e];
        __syncthreads();
        auto loss_check = [&](const DevOp& op) {
#pragma unroll
          for (int j = 0; j < NP; j++) {
            if (__builtin_expect(P.L[j].state != BMX_ALIVE, 0)) {
              if (P.live[j]) {
                P.live[j] = false;
                P.ie_lost[j] = op.ie;
                P.lost_turn[j] = t;
                on_loss(j, P.L[j], t, op.ie);
              }
              park(P.L[j], lc);
            }
          }
        };
        if (P.any()) {
          for (int k = 0; k < m;) {
            const DevOp& op = sops[k];
            int flags = __builtin_amdgcn_readfirstlane(op.flags);
#ifdef BMX_CELL_LEN
            if (__builtin_expect((flags & F_CELL) && k + BMX_CELL_LEN <= m, 1)) {
              static_assert(NP == 1, "cell kernels carry one particle per lane");
              jit_cell(P.L[0], sops + k, lc, tt, loss_check);
              k += BMX_CELL_LEN;
              continue;
            }
#endif
            int type = __builtin_amdgcn_readfirstlane(op.type);
            int n = __builtin_amdgcn_readfirstlane(op.n);
#pragma unroll
            for (int j = 0; j < NP; j++) run_op(P.L[j], op, type, flags, n, lc, tt);
            loss_check(op);
            k++;
          }
        }
        c += m;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
struct SoA {
  double *x, *px, *y, *py, *z, *pz, *beta;
  int32_t *state, *ix_lost, *turn_lost;
  double* t;  // particle time (bunch kernel, may be null)
};

// Waves per SIMD the register allocation targets (occupancy); the lane
// program is FP64-latency bound, so occupancy is traded against spills.
#ifndef BMX_WAVES_PER_EU
#define BMX_WAVES_PER_EU (BMX_NP == 1 ? 4 : 2)
#endif

// NP particles of a bunch per lane: particle j of lane `tid` in workgroup
// `blk` is base + j * 256 + tid (coalesced per j).  Load, `n_turns` turns
// over [lo1, hi1) then [lo2, hi2), store.  Particles already lost on entry are
// left untouched; a particle lost on the way is stored at its loss.  With a
// time array, a tracked particle's t is replaced by its offset from the
// reference relation, dt = t - t_start + z / (beta c), for the length of the
// walk (so that kernels without the lane's dt keep it in memory), and set to
// the reference time of its last edge + dt - z / (beta c) at the end.
template <int NP>
__device__ __forceinline__ void track_lanes(int64_t base, int64_t end, const DevOp* __restrict__ ops, int lo1,
                                            int hi1, int lo2, int hi2, const LatticeConsts& lc, int n_turns, SoA& p,
                                            unsigned long long* __restrict__ n_lost, DevOp* sops) {
  Pack<NP> P;
  int64_t idx[NP];
  bool entered[NP];
#pragma unroll
  for (int j = 0; j < NP; j++) {
    int64_t i = base + (int64_t)j * blockDim.x + threadIdx.x;
    idx[j] = i;
    entered[j] = i < end && p.state[i] == BMX_ALIVE;
    P.live[j] = false;
    P.ie_lost[j] = -1;
    P.lost_turn[j] = 0;
    park(P.L[j], lc);
    if (entered[j]) {
      Lane& L = P.L[j];
      L.x = p.x[i]; L.px = p.px[i]; L.y = p.y[i]; L.py = p.py[i]; L.z = p.z[i]; L.pz = p.pz[i]; L.beta = p.beta[i];
#if BMX_RAD
      L.pid = (uint32_t)i;
#endif
      refresh_invariants(L, lc);
      P.live[j] = !too_large(L, lc);
      if (p.t && P.live[j]) {
        double dt = p.t[i] - lc.t_start + L.z / (L.beta * D_C_LIGHT);
        p.t[i] = dt;
#if BMX_ABS_TIME
        L.dt = dt;
#endif
      }
      if (!P.live[j]) {  // lost before the first element
        p.x[i] = L.x; p.px[i] = L.px; p.y[i] = L.y; p.py[i] = L.py; p.z[i] = L.z; p.pz[i] = L.pz; p.beta[i] = L.beta;
        p.state[i] = L.state;
        if (p.ix_lost) p.ix_lost[i] = ops[lo1].ie;
        if (p.turn_lost) p.turn_lost[i] = 0;
        atomicAdd(n_lost, 1ull);
        park(L, lc);
      }
    }
  }
  walk<NP, OPS_CHUNK>(P, ops, lo1, hi1, lo2, hi2, n_turns, 0, lc, sops, [&](int j, const Lane& L, int t, int ie) {
    int64_t i = idx[j];
    p.x[i] = L.x; p.px[i] = L.px; p.y[i] = L.y; p.py[i] = L.py; p.z[i] = L.z; p.pz[i] = L.pz; p.beta[i] = L.beta;
    if (p.t) p.t[i] = edge_time(lc, ie, t) + p.t[i] - L.z / (L.beta * D_C_LIGHT);
    p.state[i] = L.state;
    if (p.ix_lost) p.ix_lost[i] = ie;
    if (p.turn_lost) p.turn_lost[i] = t;
    atomicAdd(n_lost, 1ull);
  });
#pragma unroll
  for (int j = 0; j < NP; j++) {
    if (!P.live[j]) continue;
    int64_t i = idx[j];
    const Lane& L = P.L[j];
    p.x[i] = L.x; p.px[i] = L.px; p.y[i] = L.y; p.py[i] = L.py; p.z[i] = L.z; p.pz[i] = L.pz; p.beta[i] = L.beta;
    if (p.t) p.t[i] = lc.t_end + p.t[i] - L.z / (L.beta * D_C_LIGHT);
  }
}

// Multi-turn bunch tracking: segments [lo1, hi1) then [lo2, hi2) per turn.
extern "C" __global__ __launch_bounds__(256, BMX_WAVES_PER_EU) void bmx_track_bunch_kernel(
    const DevOp* __restrict__ ops, int lo1, int hi1, int lo2, int hi2, LatticeConsts lc, int n_turns, SoA p,
    int64_t n, unsigned long long* __restrict__ n_lost) {
  __shared__ DevOp sops[OPS_CHUNK];
  int64_t base = (int64_t)blockIdx.x * blockDim.x * BMX_NP;
  track_lanes<BMX_NP>(base, n, ops, lo1, hi1, lo2, hi2, lc, n_turns, p, n_lost, sops);
}

// Particles of several variants in one launch: particles [off[v], off[v+1])
// are tracked through variant v (blockIdx.y = v, whole turns).
extern "C" __global__ __launch_bounds__(256, BMX_WAVES_PER_EU) void bmx_track_multi_kernel(
    const DevOp* __restrict__ ops, const int32_t* __restrict__ var_op_begin, const int64_t* __restrict__ off,
    LatticeConsts lc, int n_turns, SoA p, unsigned long long* __restrict__ n_lost) {
  __shared__ DevOp sops[OPS_CHUNK];
  int v = blockIdx.y;
  int64_t base = off[v] + (int64_t)blockIdx.x * blockDim.x * BMX_NP;
  track_lanes<BMX_NP>(base, off[v + 1], ops, var_op_begin[v], var_op_begin[v + 1], 0, 0, lc, n_turns, p, n_lost,
                      sops);
}

// Batched grid tracking in turn blocks with lane compaction.  Particles of
// variant v occupy slots [v * n_ic, v * n_ic + count[v]) of a structure-of-
// arrays store; blockIdx.y = variant, so every wave walks one variant's
// program.  A block tracks up to `turns` turns; survivors are appended to the
// destination store through one atomic per wave (ballot + lane prefix), lost
// particles write their survival turn.  The next block launches only over the
// survivors, so waves stay full as the grid loses particles.
struct GridDev {
  int n_angle, n_amp, n_pz, n_ic;
  double amp_max_x, amp_ratio_y, pz_min, pz_max, angle_min, angle_max;
};

struct GridSoA {
  double *x, *px, *y, *py, *z, *pz, *beta;
  int32_t* ic;
};

__global__ __launch_bounds__(256) void bmx_grid_init_kernel(GridDev g, LatticeConsts lc, GridSoA s,
                                                           const double* __restrict__ ref_orbit,
                                                           int32_t* __restrict__ survival, int n_turns) {
  int v = blockIdx.y;
  int ic = blockIdx.x * blockDim.x + threadIdx.x;
  if (ic >= g.n_ic) return;
  int i_amp = ic % g.n_amp;
  int i_ang = (ic / g.n_amp) % g.n_angle;
  int i_pz = ic / (g.n_amp * g.n_angle);
  double amp = g.amp_max_x * (double)(i_amp + 1) / (double)g.n_amp;
  double th = (g.n_angle > 1) ? g.angle_min + (g.angle_max - g.angle_min) * (double)i_ang / (double)(g.n_angle - 1)
                              : g.angle_min;
  double pz0 = (g.n_pz > 1) ? g.pz_min + (g.pz_max - g.pz_min) * (double)i_pz / (double)(g.n_pz - 1) : g.pz_min;
  double r[6] = {0.0, 0.0, 0.0, 0.0, 0.0, pz0};
  if (ref_orbit) {  // offsets are taken from the reference orbit of this variant and pz slice
    const double* o = ref_orbit + ((int64_t)v * g.n_pz + i_pz) * 6;
    for (int k = 0; k < 6; k++) r[k] = o[k];
  }
  int64_t k = (int64_t)v * g.n_ic + ic;
  s.x[k] = r[0] + amp * cos(th);
  s.y[k] = r[2] + amp * g.amp_ratio_y * sin(th);
  s.px[k] = r[1]; s.py[k] = r[3]; s.z[k] = r[4]; s.pz[k] = r[5];
  double pc = lc.p0c * (1.0 + r[5]);
  s.beta[k] = (r[5] == 0.0) ? lc.beta0 : pc / sqrt(pc * pc + lc.mc2sq);
  s.ic[k] = ic;
  survival[k] = n_turns;
}

// Explicit initial conditions already placed in slots [v * stride, v * stride + count[v]):
// beta from pz (init_coord), slot ids, survival = n_turns.
__global__ __launch_bounds__(256) void bmx_ics_init_kernel(LatticeConsts lc, GridSoA s, const int32_t* __restrict__ count,
                                                          int stride, int32_t* __restrict__ survival, int n_turns) {
  int v = blockIdx.y;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count[v]) return;
  int64_t k = (int64_t)v * stride + i;
  double pz = s.pz[k];
  double pc = lc.p0c * (1.0 + pz);
  s.beta[k] = (pz == 0.0) ? lc.beta0 : pc / sqrt(pc * pc + lc.mc2sq);
  s.ic[k] = i;
  survival[k] = n_turns;
}

extern "C" __global__ __launch_bounds__(GRID_BLOCK, BMX_WAVES_PER_EU) void bmx_grid_block_kernel(
    const DevOp* __restrict__ ops, const int32_t* __restrict__ var_op_begin, const int32_t* __restrict__ var_n_ele,
    LatticeConsts lc, int n_ic, GridSoA src, GridSoA dst, const int32_t* __restrict__ count_in,
    int32_t* __restrict__ count_out, const int32_t* __restrict__ d_turn0, int turns_per_block, int n_turns,
    int32_t* __restrict__ survival, unsigned long long* __restrict__ pet, const int32_t* __restrict__ var_order) {
  __shared__ DevOp sops[GRID_BLOCK];
  // the block's first turn lives in device memory so that captured graphs of
  // several blocks replay unchanged; every thread reads the same value, so a
  // finished scan leaves the whole grid together
  const int turn0 = *d_turn0;
  const int turns = min(turns_per_block, n_turns - turn0);
  if (turns <= 0) return;
  constexpr int NP = BMX_NP;
  // the variants in dispatch order, longest program first: the workgroups
  // are handed out in blockIdx order, so the long ones start first and the
  // short ones fill the compute units they leave (longest-processing-time
  // order)
  int v = var_order[blockIdx.y];
  int64_t base = (int64_t)v * n_ic;
  int n_in = count_in[v];
  Pack<NP> P;
  int slot[NP], ic[NP];
  bool active[NP];
#pragma unroll
  for (int j = 0; j < NP; j++) {
    slot[j] = (blockIdx.x * NP + j) * blockDim.x + threadIdx.x;
    active[j] = slot[j] < n_in;
    ic[j] = 0;
    P.live[j] = false;
    P.ie_lost[j] = -1;
    P.lost_turn[j] = 0;
    park(P.L[j], lc);
    if (active[j]) {
      int64_t k = base + slot[j];
      Lane& L = P.L[j];
      L.x = src.x[k]; L.px = src.px[k]; L.y = src.y[k]; L.py = src.py[k]; L.z = src.z[k]; L.pz = src.pz[k];
      L.beta = src.beta[k];
      ic[j] = src.ic[k];
      refresh_invariants(L, lc);
      P.live[j] = !(turn0 == 0 && too_large(L, lc));
      if (!P.live[j]) {
        survival[base + ic[j]] = turn0;
        park(L, lc);
      }
    }
  }
  int lo = var_op_begin[v], hi = var_op_begin[v + 1];
  walk<NP, GRID_BLOCK>(P, ops, lo, hi, 0, 0, turns, turn0, lc, sops,
           [&](int j, const Lane&, int t, int) { survival[base + ic[j]] = turn0 + t; });
  // live particle-element-turns of this lane
  unsigned long long my_pet = 0;
#pragma unroll
  for (int j = 0; j < NP; j++)
    if (active[j]) {
      unsigned long long tr = P.live[j] ? (unsigned long long)turns : (unsigned long long)P.lost_turn[j];
      my_pet += tr * (unsigned long long)var_n_ele[v] + (P.live[j] ? 0ull : (unsigned long long)(P.ie_lost[j] + 1));
    }
  // wave-level compaction of survivors, one ballot per particle slot
  int lane = __lane_id();
#pragma unroll
  for (int j = 0; j < NP; j++) {
    bool keep = active[j] && P.live[j];
    unsigned long long mask = __ballot(keep);
    unsigned long long below = (lane == 0) ? 0ull : (mask & ((~0ull) >> (64 - lane)));
    int n_alive = __popcll(mask);
    int first = __ffsll((long long)mask) - 1;
    int slot0 = 0;
    if (n_alive > 0 && lane == first) slot0 = atomicAdd(&count_out[v], n_alive);
    slot0 = __shfl(slot0, first < 0 ? 0 : first);
    if (keep) {
      int64_t k = base + slot0 + __popcll(below);
      const Lane& L = P.L[j];
      dst.x[k] = L.x; dst.px[k] = L.px; dst.y[k] = L.y; dst.py[k] = L.py; dst.z[k] = L.z; dst.pz[k] = L.pz;
      dst.beta[k] = L.beta;
      dst.ic[k] = ic[j];
    }
  }
  for (int off = 32; off > 0; off >>= 1) my_pet += __shfl_xor(my_pet, off);
  if (lane == 0 && my_pet) atomicAdd(pet, my_pet);
}

// Step bookkeeping between turn blocks: clears the count buffer the next
// block appends survivors to and advances the block's first turn.  A kernel,
// not a memset, so that it can sit in a captured graph.
__global__ __launch_bounds__(256) void bmx_grid_advance_kernel(int32_t* __restrict__ count_next, int nv,
                                                              int32_t* __restrict__ d_turn0, int turns) {
  for (int v = threadIdx.x; v < nv; v += blockDim.x) count_next[v] = 0;
  if (threadIdx.x == 0) *d_turn0 += turns;
}

// Bunch moments over live particles (the sums of ltt_calc_bunch_data,
// bsim/modules/lt_tracking_mod.f90:2391-2402): out[0] = n_live,
// out[1..6] = sum vec_i, out[7..27] = sum vec_i vec_j (i <= j, row-major upper
// triangle).  Two passes with a fixed reduction order, so the sums do not
// depend on scheduling: per-block partials, then one block over the partials.
constexpr int N_MOM = 28;

__device__ __forceinline__ void block_reduce_moments(double (&a)[N_MOM], double* red) {
  // red: N_MOM x 256 doubles of LDS
  int t = threadIdx.x;
  for (int k = 0; k < N_MOM; k++) red[k * 256 + t] = a[k];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w)
      for (int k = 0; k < N_MOM; k++) red[k * 256 + t] += red[k * 256 + t + w];
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) void bmx_moments_partial_kernel(SoA p, int64_t n, double* __restrict__ partial) {
  extern __shared__ double red[];
  double a[N_MOM];
  for (int k = 0; k < N_MOM; k++) a[k] = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    if (p.state[i] != BMX_ALIVE) continue;
    double v[6] = {p.x[i], p.px[i], p.y[i], p.py[i], p.z[i], p.pz[i]};
    a[0] += 1.0;
    int k = 7;
    for (int r = 0; r < 6; r++) {
      a[1 + r] += v[r];
      for (int c = r; c < 6; c++) a[k++] += v[r] * v[c];
    }
  }
  block_reduce_moments(a, red);
  if (threadIdx.x < N_MOM) partial[(int64_t)blockIdx.x * N_MOM + threadIdx.x] = red[threadIdx.x * 256];
}

__global__ __launch_bounds__(256) void bmx_moments_final_kernel(const double* __restrict__ partial, int n_part,
                                                               double* __restrict__ out) {
  extern __shared__ double red[];
  double a[N_MOM];
  for (int k = 0; k < N_MOM; k++) a[k] = 0.0;
  for (int b = threadIdx.x; b < n_part; b += 256)
    for (int k = 0; k < N_MOM; k++) a[k] += partial[(int64_t)b * N_MOM + k];
  block_reduce_moments(a, red);
  if (threadIdx.x < N_MOM) out[threadIdx.x] = red[threadIdx.x * 256];
}
