// Host-side context of the liquid-substep library and the launchers each
// kernel module exports to the C ABI layer (wc_capi.cpp).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/wetcloth_amd.h"
#include "wc_device.h"

namespace wc {

// A HIP error inside a launcher is turned into this and caught at the ABI.
struct HipError {
  hipError_t code;
  const char* what;
  const char* file;
  int line;
};

#define WC_HIP(expr)                                         \
  do {                                                       \
    hipError_t _e = (expr);                                  \
    if (_e != hipSuccess) throw ::wc::HipError{_e, #expr, __FILE__, __LINE__}; \
  } while (0)
#define WC_LAUNCH_CHECK() WC_HIP(hipGetLastError())

template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t cap = 0;  // elements
  // A buffer that has to grow again (the DOF count and the sweep plan vary
  // from substep to substep) gets 25 % headroom, so that the free and
  // allocation, which stall the device, are rare.
  bool lent = false;  // memory lent by the caller: never freed or regrown here
  void lend(T* q, size_t n) {
    release();
    p = q;
    cap = n;
    lent = true;
  }
  void ensure(size_t n) {
    if (n <= cap && p) return;
    if (lent)
      throw HipError{hipErrorOutOfMemory, "a lent buffer is smaller than the request", __FILE__,
                     __LINE__};
    const bool regrow = p != nullptr;
    if (p) WC_HIP(hipFree(p));
    p = nullptr;
    cap = 0;
    size_t c = n ? n : 1;
    if (regrow) c += c / 4;
    WC_HIP(hipMalloc((void**)&p, c * sizeof(T)));
    cap = c;
  }
  void release() {
    if (p && !lent) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    lent = false;
  }
};

// Matrix in "padded CSR": row i holds len[i] sorted entries starting at
// start[i]; with start == nullptr rows start at i*stride.
struct DevCsr {
  int64_t n = 0;
  const uint32_t* start = nullptr;
  const uint32_t* len = nullptr;
  const uint32_t* rowptr = nullptr;  // compact CSR (start[i] .. rowptr[i+1])
  uint32_t stride = 0;
  const uint32_t* col = nullptr;
  const double* val = nullptr;
  int64_t nnz_slots = 0;  // size of col/val arrays
};

// Optional timing of the solver's kernels with HIP events recorded on the
// launching stream around each launch; pending pairs are read back only at
// host synchronization points the solver already has.
struct KTimer {
  enum Kind { SWEEP_FWD = 0, SWEEP_BWD = 1, SPMV = 2, P2G = 3, G2P = 4, COUNT = 5 };
  bool on = false;
  int stride = 1;    // time the launches of every stride-th PCG iteration
  int64_t tick = 0;
  // this timer if the next PCG iteration is one to time, else null
  KTimer* sample() { return on && tick++ % stride == 0 ? this : nullptr; }
  double ms[COUNT] = {};
  int64_t launches[COUNT] = {};
  int64_t units[COUNT] = {};
  struct Rec {
    int kind;
    int64_t units;
    hipEvent_t a, b;
  };
  std::vector<Rec> pending;
  std::vector<hipEvent_t> spare;
  hipEvent_t mark(hipStream_t st);  // record a fresh event on st
  void add(int kind, int64_t units, hipEvent_t a, hipEvent_t b) {
    pending.push_back(Rec{kind, units, a, b});
  }
  void harvest();  // the stream has been synchronized: accumulate pending
  void harvest_ready();  // accumulate the pending pairs that have completed
  void reset();
  void release();
};

// The pipelined sweeps' copy of a 7-point system (wc_sweep.hip): the DOF
// bounding box, cut into 16x16 tiles of k-pencils and stored in the skewed
// tile layout, where the cells a sweep step touches are consecutive.
struct SweepPlan {
  bool valid = false;
  int o[3] = {0, 0, 0}, B[3] = {0, 0, 0};
  int ntile_a = 0, ntile_b = 0, ntiles = 0;
  int nsteps = 0;       // skewed steps per tile (B[2] + 30)
  int64_t ncell = 0;    // slots of the skewed tile layout
  int64_t kdx = 0, kdy = 0;  // slot offsets to the upstream x / y face cell
  DevBuf<int32_t> pdof;             // box cell -> DOF or -1
  DevBuf<uint32_t> order, code;     // ticket order; per-cell coupling bits
  DevBuf<uint32_t> scode;           // per-cell sweep code (orders applied)
  DevBuf<uint32_t> opf, opb;        // per-cell LDS operand offsets of each sweep
  DevBuf<uint16_t> sflag;           // per-cell SpMV flags (M_* in wc_sweep.hip)
  DevBuf<double> cf, cb;            // 32-byte factor records (forward / backward)
  DevBuf<double> coef;              // 7 matrix entries per cell
  DevBuf<double> fin;               // 96-byte factor-pass inputs per cell
  DevBuf<double> rsum;              // 80-byte row summaries per cell (plan fill only)
  DevBuf<double> zf, zb;            // sweep outputs (zb = preconditioned vector)
  DevBuf<double> rb, sb, xb;        // PCG vectors r, s, x in box layout
  // mixed-precision solve: FP32 copies of the records and entries, FP32
  // sweep outputs and PCG vectors, and b in box layout (FP64)
  DevBuf<float> cf32, cb32, coef32, zf32, zb32, rb32, sb32;
  DevBuf<double> bb;
  DevBuf<int> progress, flags;
  DevBuf<double> tdot;              // per-unit z . r partials of the last apply
  int ndot = 0;                     // how many (tiles, or sub-tiles)
  // Sub-tile sweeps (one wave per 8x8 quarter of a tile, k_sweep_sub) or
  // the tile sweeps (k_sweep); both give the same values bit for bit
  bool sub = true;
  DevBuf<uint32_t> order_sub;       // sub-tile ticket order
  // The windowed SpMV (k_box_spmv_win: one block per tile and run of
  // spmv_run steps) or the slot-parallel one; both give the same y bit for
  // bit.  spart: the windowed product's x . y partials, one per block.
  bool debug_abort = false;  // test hook: the sweeps' abort flag set before each PCG loop
  bool spmv_win = true;
  int spmv_run = 12;
  DevBuf<double> spart;
  // its x faces as compact columns (ncell / 16 each): li = 15 and li = 0 of
  // the operand (written by the s update, else extracted), and li = 15 of
  // the -x coupling plane (once per plan)
  DevBuf<double> cf15, xf15, xf0;
  DevBuf<float> cf15_32, xf15_32, xf0_32;
  // The forward sweep's ticket counter is zero: the sweeps zero each
  // other's counter, so only the first apply after the plan needs a clear.
  bool ticks_zero = false;
  void release();
};

// Where the rows of a structured (grid) system live, so the solver can build
// the pipelined sweeps; all null for a general CSR matrix.
struct GridSystem {
  const Grid* g = nullptr;
  const uint8_t* act = nullptr;
  const uint32_t* isdof = nullptr;
  const uint32_t* dofscan = nullptr;
  const uint32_t* rowlen = nullptr;
  const uint32_t* col = nullptr;
};

// Device work space of the MIC(0)-PCG solver (PCGSolver.h:310-314).
struct PcgWork {
  GridSystem grid;   // set by the caller for grid systems
  KTimer timer;
  bool factored = false;  // a factor of the last system exists (either path)
  bool use_sweeps = true;
  // mixed-precision solve (grid systems with the pipelined sweeps)
  bool mixed = false;
  int mixed_digits = 6;       // residual reduction per refinement pass, 10^-digits
  int mixed_refinements = 2;  // refinement passes after the first
  int passes = 0;             // passes of the last mixed solve
  SweepPlan sweep;
  DevBuf<double> invdiag, adiag, lv, r, z, s, x, partial, scal;
  DevBuf<uint32_t> tpos;
  DevBuf<int> flags;  // [0] relaxation changed, [1] factor overflow, [2] max level
  DevBuf<int32_t> level;
  DevBuf<uint32_t> lvl_count, lvl_cursor, lvl_list;
  std::vector<uint32_t> lvl_start;  // host copy: start of each level
  std::vector<uint32_t> lvl_size;
  int nlevels = 0;
  const int32_t* level_in = nullptr;  // caller's per-row level, or null
  int32_t level_span = 0;
  bool levels_ready = false;
  int red_blocks = 0;
  double* host_scal = nullptr;  // pinned, or host_scal_own when no_alloc
  bool no_alloc = false;        // every buffer lent (DevBuf::lend); no pinned memory either
  double host_scal_own[8] = {};
  double* host_part = nullptr;  // mapped host memory, red_blocks |r| partials
  double* host_part_dev = nullptr;  // its device address
  void release();
};

// The AMG preconditioner of AMGPCGSolveSparse (AlgebraicMultigrid.h,
// GeometricLevelGen.h) on the device (wc_amg.hip).
struct AmgLevel {
  int64_t n = 0;
  int ni = 0, nj = 0, nk = 0;   // grid of the DOFs' cell coordinates
  DevCsr A;                     // level 0: the caller's rows; coarser: 7 slots per row
  DevBuf<uint32_t> col, len;    // coarse rows (7 slots each, len used)
  DevBuf<double> val;
  DevBuf<int32_t> ijk;          // 3 per DOF
  DevBuf<uint8_t> red;          // (i + j + k) odd: smoothed first
  DevBuf<uint32_t> parent;      // the coarse DOF of each DOF (every level but the last)
  DevBuf<uint32_t> kids;        // 8 per DOF: its fine children ascending, ~0u after them
                                // (every level but the first)
  DevBuf<double> x, b, r;       // V-cycle vectors (x, b of level 0 are the caller's)
  // the rows again, ordered by colour (red rows ascending, then black rows
  // ascending) with each row's entries in their order as 7 planes of n, so a
  // colour's smoothing reads only its own rows, coalesced
  int64_t nred = 0;
  DevBuf<uint32_t> crow;        // compacted position -> row
  DevBuf<uint32_t> ccol;        // [s * n + p]
  DevBuf<double> cval;
  DevBuf<uint8_t> clen;
};
constexpr int kAmgMaxLevels = 16;
struct AmgHier {
  AmgLevel L[kAmgMaxLevels];
  int nl = 0;
  bool compact = false;         // every level has its colour-ordered rows
  DevBuf<uint32_t> cmin, flag, scan_tmp;
  DevBuf<uint32_t> nred;        // red rows of each level, on the device
  DevBuf<int> err;
  void release();
};
// builds the levels of A (level 0 rows as given, DOF cell coordinates ijk
// on the device, 3 per DOF, in an ni x nj x nk grid)
void amg_build(hipStream_t st, AmgHier& H, const DevCsr& A, const int32_t* ijk, int ni,
               int nj, int nk);
// z = one V-cycle applied to b from z = 0 (amgPrecondCompressed)
void amg_vcycle(hipStream_t st, AmgHier& H, const double* b, double* z);
// y = A x on level 0's colour-ordered rows (the products added in the row's
// order from 0, as multiply); false when the hierarchy has none
bool amg_spmv(hipStream_t st, AmgHier& H, const double* x, double* y);
// AMGPCGSolveSparse (AlgebraicMultigrid.h:247-359) on a built hierarchy
bool amgpcg_solve(hipStream_t st, PcgWork& w, AmgHier& H, const DevCsr& A, const double* b,
                  double* x, double tol, int maxit, double* residual_out, int* iters_out);

struct PhaseEvents {
  hipEvent_t ev[8] = {};
  bool have = false;
};

// The adjoint (wc_adjoint.hip; the reference has no counterpart).  The
// decisions of one recorded substep (FrozenView): act per bucket; per node
// liquid, flags, and arg = the caller index of the particle whose distance
// gives the cell's liquid phi (-1: the default three buckets, -2: the solid
// override); proj per caller index.
struct FrozenRecord {
  DevBuf<uint8_t> act, liquid, flags, proj;
  DevBuf<int32_t> arg;
  bool recorded = false;
  int64_t np = 0;  // particles of the recorded substep
  void release();
};
// a particle state in the caller's order (x[d*n+p], C[t*n+p])
struct ParticleCheckpoint {
  DevBuf<double> x, v, C;
  int64_t n = 0;
  void release();
};
// device state of the adjoint substep
struct AdjointWork {
  DevBuf<uint32_t> slot_of, argmin;       // caller index -> slot; per node scratch
  DevBuf<double> x0, v0, C0;              // the substep's start state, slot order
  DevBuf<double> gx, gv, gC;              // dL/d(particle state), caller order
  DevBuf<double> ax;                      // dL/dx0 of the G2P part, slot order
  DevBuf<double> ub[3], ub2[3], ub0[3];   // face gradients
  DevBuf<uint8_t> code[3], cnt[3];        // extrapolation layers per face
  DevBuf<double> phib, pb, bb, partial;   // per cell, per DOF, reduction partials
  int64_t n = 0;                          // particles of the gradient state
  bool have_start = false;                // x0.. belong to the last substep
  bool ready = false;                     // the last call ran a whole substep in mode 1 or 2
  void release();
};

// The static porous solid (wc_mix.hip): its particles, LiquidInfo's
// mixture fields, the solid's face sums (taken once: the solid rests) and
// the per-substep mixture fields.
struct MixState {
  int64_t ne = 0;  // solid particles; 0 = liquid only
  wc_mix_params prm{};
  DevBuf<double> x, m, fm, fvol, rvol, rvf, ori, sf;  // x, ori: 3 planes of ne
  // per face axis d, 8 planes of nnode: solid mass, solid volume (rest_vol
  // rest_vf), carried liquid volume, shape factor, solid volume incl. the
  // carried liquid, orientation x, y, z (plane())
  DevBuf<double> acc;
  DevBuf<double> cvs, cvf, cvl;  // cell sums: solid volume, its liquid, the liquid particles'
  DevBuf<uint8_t> touch, near;   // per bucket: touched by the solid's stencils; next to one
  DevBuf<double> psi[3], sat[3], Q[3], hdv[3];
  DevBuf<double> psi_p, sat_p, pp;  // per cell
  double* plane(const Grid& g, int d, int k) const;
  void release();
};

}  // namespace wc

struct wc_ctx {
  wc_desc desc;
  wc::Grid g;
  hipStream_t stream = nullptr;
  std::string err;

  // particles, SoA: x[d*cap+p], v[d*cap+p], C[(r*3+c)*cap+p]; pid = the
  // caller's particle index (the device keeps particles sorted by bucket).
  // FP64 planes for WC_FP64, FP32 planes (fx, ...) for WC_MIXED; the other
  // set stays unallocated.
  int64_t np = 0, pcap = 0;
  bool mixed = false;
  wc::DevBuf<double> x, v, C, x2, v2, C2;
  wc::DevBuf<float> fx, fv, fC, fx2, fv2, fC2;
  wc::DevBuf<uint32_t> pid, pid2, pkey;
  wc::DevBuf<uint32_t> bcount, bstart, bcursor;  // per bucket (+1)
  bool sorted = false;  // bstart describes the current particle arrays
  // P2G sums in fixed point (order independent), so that a run repeats
  // bit for bit (wc_set_option "deterministic")
  bool deterministic = false;
  // positions before the last separate advection pass (wc_advect_scene), the
  // ones the transfer weights belong to; solidProjection weighs at them
  wc::DevBuf<double> xw;
  wc::DevBuf<float> fxw;
  bool xw_valid = false;
  wc::DevBuf<unsigned long long> vmax_bits;

  // grid (nnode entries unless noted)
  wc::DevBuf<uint8_t> touched, act;  // per bucket
  wc::DevBuf<double> mass[3], mom[3], vol[3];  // mom holds velocity after P2G
  wc::DevBuf<unsigned long long> phikey;
  wc::DevBuf<double> lphi, w[3], rhs[3], invC[3], ustar[3], unew[3], utmp[3];
  wc::DevBuf<uint8_t> valid[3], vtmp[3];
  wc::DevBuf<uint32_t> isdof, dofscan;  // nnode + 1
  wc::DevBuf<double> pressure;

  // assembled system: ELL-7 rows in DOF order
  int64_t ndof = 0;
  wc::DevBuf<uint32_t> rowlen, col;
  wc::DevBuf<double> val, b, p;
  wc::DevBuf<int32_t> dof_level;
  int32_t level_min = 0, level_max = 0;
  wc::PcgWork pcg;

  // scan scratch
  wc::DevBuf<uint32_t> scan_tmp;

  double last_residual = 0.0;
  int last_iters = 0;
  int last_substeps = 0;  // substeps of the last wc_step (the CFL split)
  // pressure solver: 0 = MIC(0) PCGSolver, 1 = AMGPCGSolveSparse (the
  // reference's call at Pressure.cpp:797), with its hierarchy and the DOFs'
  // cell coordinates
  int solver = 0;
  wc::AmgHier amg;
  wc::DevBuf<int32_t> dof_ijk;
  double phase_ms[7] = {0, 0, 0, 0, 0, 0, 0};
  wc::PhaseEvents events;

  // the adjoint: decisions recorded (frz_mode 1) into or replayed (2) from
  // frz[frz_slot]; particle checkpoints; the adjoint substep's state; the
  // last rollout's length and checkpoint spacing
  int frz_mode = 0, frz_slot = 0;
  std::vector<wc::FrozenRecord> frz;
  wc::DevBuf<unsigned long long> frz_dep;  // departures of the replayed substeps
  std::vector<wc::ParticleCheckpoint> ckpt;
  wc::AdjointWork adj;
  int roll_steps = 0, roll_every = 0;

  // a static porous solid in the liquid path (wc_set_elasto_particles)
  wc::MixState mix;
};

namespace wc {

// ---- wc_scan.hip ----
// In-place exclusive scan of n uint32 values; returns the total.
uint32_t exclusive_scan(wc_ctx* ctx, uint32_t* data, int64_t n);
uint32_t exclusive_scan(hipStream_t st, DevBuf<uint32_t>& scratch, uint32_t* data, int64_t n);
// the same without reading the total back (no host wait); the scratch must
// not be reused before the scan has run
void exclusive_scan_async(hipStream_t st, DevBuf<uint32_t>& scratch, uint32_t* data, int64_t n);

// ---- wc_particles.hip ----
void particles_sort(wc_ctx* ctx);
void particles_p2g(wc_ctx* ctx);
void particles_g2p_advect(wc_ctx* ctx, double dt);
// the three parts of particles_g2p_advect as separate passes:
// mapNodeParticlesAPIC, advectScene, solidProjection
void particles_g2p(wc_ctx* ctx);
void particles_advect(wc_ctx* ctx, double dt);
void particles_solid_projection(wc_ctx* ctx, double dt);
// getMaxFluidVelocity (waits for the stream)
double particles_max_velocity(wc_ctx* ctx);
// the particles' bounding box (waits for the stream)
void particles_bbox(wc_ctx* ctx, double lo[3], double hi[3]);
int64_t particles_seed_block(wc_ctx* ctx, const double lo[3], const double hi[3],
                             uint64_t seed, double sigma);
void particles_set_soa(wc_ctx* ctx, int64_t n, const double* x, const double* v,
                       const double* C, bool aos_ref_layout);
void particles_get_soa(wc_ctx* ctx, double* x, double* v, double* C,
                       bool aos_ref_layout);

// ---- wc_adjoint.hip ----
// the decisions of the current mode and record (mode 0: all null)
FrozenView frozen_view(wc_ctx* ctx);
// the record of the current slot, sized for the grid and particles
FrozenRecord& frozen_record(wc_ctx* ctx);
// mode 2: the liquid phi of every cell from its recorded particle
void frozen_replay_lphi(wc_ctx* ctx);
// mode 1: the particle whose distance gave each cell's liquid phi
void frozen_record_lphi_arg(wc_ctx* ctx);
// modes 1 and 2, after the substep's sort: the slot of every caller index and
// a copy of the start state
void frozen_after_sort(wc_ctx* ctx);
// the adjoint of the substep just run (mode 1 or 2) on the device gradient
// state: gx, gv, gC go from dL/d(state after) to dL/d(state before); *gb
// accumulates dL/d|g| for gravity (0, -|g|, 0).  Returns whether the
// adjoint pressure solve converged.
bool adjoint_substep(wc_ctx* ctx, double dt, double tol, int maxit, double* gb,
                     double* residual_out, int* iters_out);
// device copies of the particle state in the caller's order (FP64 contexts)
void checkpoint_save(wc_ctx* ctx, int slot);
void checkpoint_restore(wc_ctx* ctx, int slot);
void adjoint_release(wc_ctx* ctx);

// ---- wc_mix.hip ----
// the solid's particles and parameters; its face and cell sums (ne = 0
// removes the solid)
void mix_set(wc_ctx* ctx, int64_t ne, const double* x, const double* m, const double* fm,
             const double* fvol, const double* rvol, const double* rvf, const double* ori,
             const double* sf, const wc_mix_params& prm);
// after the liquid's P2G: the buckets the solid's stencils touch
void mix_touch(wc_ctx* ctx);
// after grid_nodes: saturations, the cell-centred fields and the pore pressure
void mix_grid(wc_ctx* ctx);
// the grid update with the pore pressure gradient and the drag
void mix_step_velocity(wc_ctx* ctx, double dt);
// the plane indices of MixState::acc
constexpr int kMixPlaneMass = 0, kMixPlaneSf = 3, kMixPlaneSvol = 4, kMixPlaneOri = 5;

// ---- wc_grid.hip ----
void grid_activate(wc_ctx* ctx);
void grid_nodes(wc_ctx* ctx);
void grid_step_velocity(wc_ctx* ctx, double dt);
int64_t grid_assemble(wc_ctx* ctx, double dt);
void grid_scatter_pressure(wc_ctx* ctx);
// each DOF's cell (i, j, k) minus origin[3] (dense node coordinates), 3 per
// DOF, into ijk
void grid_dof_cells(wc_ctx* ctx, int32_t* ijk, const int origin[3]);
void grid_apply_pressure(wc_ctx* ctx, double dt);

// ---- wc_sweep.hip ----
void sweep_prepare(hipStream_t st, SweepPlan& P, const Grid& g, const uint8_t* act,
                   const uint32_t* isdof, const uint32_t* dofscan,
                   const uint32_t* rowlen, const uint32_t* col, const double* val,
                   int64_t ndof);
// MIC(0) factor of the prepared box (pipelined), into the sweep records
void sweep_factor(hipStream_t st, SweepPlan& P, double tau, double mdr);
// MIC(0) apply on box vectors: r (box) -> P.zb; the backward sweep also
// leaves the per-unit partials of z . r in P.tdot (P.ndot of them)
// cleared: the forward output already holds the sentinel (the PCG update
// that precedes an apply in the loop sets it), so the apply skips its fill
void sweep_apply(hipStream_t st, SweepPlan& P, const double* r, KTimer* timer,
                 int64_t ndof, bool cleared = false);
void box_spmv(hipStream_t st, SweepPlan& P, const double* x, double* y);
// mixed-precision solve: FP32 records/entries from the FP64 factor and A,
// FP32 sweeps (result in P.zb32), FP32 SpMV with FP64 row sums, and the FP64
// refinement residual r = b - A x rounded to FP32 with |r|_inf partials
void sweep_narrow(hipStream_t st, SweepPlan& P);
void sweep_apply32(hipStream_t st, SweepPlan& P, const float* r, KTimer* timer,
                   int64_t ndof, bool cleared = false);
// y = A x and the partials of x . y: with the windowed product one per
// block in P.spart, else nblocks of them in `partial` (the solver's
// reduction layout).  Returns where they are and how many.
struct DotParts {
  const double* p;
  int n;
};
// faces_ready: the windowed product's x-face columns of x (P.xf15 / xf0,
// or the FP32 ones) are current, as after the s update that wrote them.
DotParts box_spmv_dot(hipStream_t st, SweepPlan& P, const double* x, double* y,
                      double* partial, int nblocks, bool faces_ready = false);
DotParts box_spmv32_dot(hipStream_t st, SweepPlan& P, const float* x, float* y,
                        double* partial, int nblocks, bool faces_ready = false);
void box_residual(hipStream_t st, const SweepPlan& P, const double* b, const double* x,
                  float* r, double* partial, int nblocks);
void box_gather(hipStream_t st, const SweepPlan& P, const double* dof, double* box);
void box_scatter(hipStream_t st, const SweepPlan& P, const double* box, double* dof);
// whether a sweep gave up a bounded wait since the last factor or check
// (the flag is cleared by reading it)
bool sweep_aborted(hipStream_t st, SweepPlan& P);

// ---- wc_pcg.hip ----
// Levels of the lower-triangular dependency DAG: given per row (structured
// grid: i+j+k) or computed by relaxation when level == nullptr.
void pcg_prepare(hipStream_t st, PcgWork& w, const DevCsr& A,
                 const int32_t* level, int32_t level_span);
// only the vectors, partials and scalars of the PCG loop (AMG-PCG, which has
// no MIC(0) factor or level schedule)
void pcg_prepare_vectors(hipStream_t st, PcgWork& w, int64_t n);
void pcg_factor(hipStream_t st, PcgWork& w, const DevCsr& A, double tau,
                double mdr);
void pcg_apply(hipStream_t st, PcgWork& w, const DevCsr& A, const double* r,
               double* z);
void pcg_spmv(hipStream_t st, const DevCsr& A, const double* x, double* y);
// PCGSolver::solve on device vectors b -> x; returns true when converged.
bool pcg_solve(hipStream_t st, PcgWork& w, const DevCsr& A, const double* b,
               double* x, double tol, int maxit, double tau, double mdr,
               double* residual_out, int* iters_out);

}  // namespace wc
