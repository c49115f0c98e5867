// context.h -- device state of one solve and the internal API that the C ABI
// (piks3d.hip) and the request server (serve.hip) share.
//
// A context is built in two phases: ctx_create_host does the host work (the
// velocity grid and the image preprocessing, no HIP call, safe on any
// thread), ctx_create_device allocates and uploads on one stream.  After that
// every operation is enqueued on the context's stream; only the functions
// that say so synchronise.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/piks3d.h"
#include "geometry.h"

namespace piks {

struct Dims {
  int Nx, Ny, Nz, Nxt;
  int64_t Nxyt, Ntotal;
  int Nc8;
  // storage of f: octant o keeps its cells in family ofam(o)'s order, pos[F]
  // maps a cell to that order (unstored cells to the zero slot); ostride =
  // elements per octant
  const int32_t* pos[4];
  int64_t ostride;
  // right-hand sides (Knudsen numbers) sharing the geometry: kernels take the
  // one of blockIdx.y; element strides between them
  int64_t rs_f, rs_fw, rs_buf, rs_snap;
};

// Moments of one cell (src/flow.F90:72 keeps Rho, Ux, Uy, Uz as four
// arrays); interleaved so that a sweep task loads them with one scalar load.
struct alignas(32) Mom {
  double r, x, y, z;
};

// Sweep task record (see Geometry::hp_task), loaded with one scalar load.
struct alignas(32) TaskRec {
  int32_t w[8];
};

// Storage window of one (hyperplane, octant) as the sweep addresses it: the
// window's first element in f and its length, both in elements of f.
struct alignas(16) SweepWin {
  int64_t off;
  uint32_t len;
  uint32_t pad;
};

// Coefficients of the update for one combination of first-order flags
// (bit 0: x, 1: y, 2: z): (C1, C2) of each axis and the axis' term of the
// denominator, as src/solver.F90:206-214 selects them.
struct alignas(16) SweepCoef {
  double x1, x2, y1, y2, z1, z2;
  double dx, dy, dz, pad;
};

// Device state that carries from one iteration to the next, in one precision.
template <typename T>
struct State {
  T* f = nullptr;        // [octant][storage index][velocity]
  T* fw = nullptr;       // [corner][octant][velocity]
  T* buf_e = nullptr;    // x = Nx plane after the previous sweep (ghost x = 0 next time)
  T* buf_w = nullptr;    // x = 1 plane after the previous sweep (ghost x = Nx+1 next time)
  T* snap = nullptr;     // walls read by earlier coupled walls, before the wall pass
};

struct Ctx {
  piks_params prm{};
  Geometry g;
  Dims d{};
  int nc_fund = 0;
  std::vector<double> cx, cy, cz, w;
  double mu = 0, diff_flux = 0, mass = 1.0;
  // several right-hand sides: Kn, mu and the stored mass of each (rhs 0 is
  // the single solve's)
  int nrhs = 1;
  std::vector<double> kn_r, mu_r, mass_r;
  double* d_mu = nullptr;
  int exact = 0;               // serial reductions in the reference's order
  hipStream_t stream = nullptr;
  bool own_stream = true;      // false: the stream belongs to the caller
  bool stream_ordered = false; // allocations with hipMallocAsync / hipFreeAsync on `stream`
  int device = 0;
  int64_t bytes = 0;
  // device state: FP64 always; FP32 too for PIKS_PREC_FP32_REFINED contexts
  State<double> s64;
  State<float> s32;
  int phase = PIKS_PREC_FP64;  // precision the next iterations run in
  size_t n_f = 0, n_fw = 0, n_buf = 0, n_snap = 0;
  Mom* mom = nullptr;
  double* part = nullptr;
  double* plane = nullptr;
  double *dcx = nullptr, *dcy = nullptr, *dcz = nullptr, *dw = nullptr;
  uint8_t* lab = nullptr;
  int32_t* wid = nullptr;
  uint64_t* views = nullptr;
  int32_t *wall_cell = nullptr, *wall_corner = nullptr;
  uint8_t* wall_label = nullptr;
  double *e1 = nullptr, *e2 = nullptr;
  int32_t *uncoupled = nullptr, *coupled = nullptr, *coupled_snap = nullptr, *snap_cell = nullptr;
  int32_t* commit_walls = nullptr;
  int32_t* wall_pos = nullptr;   // Geometry::wall_pos
  int32_t *fluid_int = nullptr, *fluid_bnd = nullptr, *fluid_all = nullptr;
  // sweep schedule: tasks of each oriented hyperplane
  bool xcd_remap = true;       // contiguous task ranges per XCD
  int sweep_block = 64;        // threads per sweep workgroup (one wave: a task's
                               // workgroup ends with its wave, measured fastest)
  bool sweep_recip = false;    // fast mode: multiply by a tabulated 1 / denominator
  bool sweep_generic = false;  // single solve through the multi-right-hand-side kernel
  int sweep_nt = 1;            // tasks per wave of the single-solve sweep (k_sweep1)
  // single-launch sweep (k_sweep_flow; FP64 iterations of a single solve with
  // 64 velocities per octant on a context with its own stream): octant o on
  // XCD o, tasks split statically over flow_waves waves per octant, or claimed
  // from a counter when the octant's waves are not all resident
  int flow_req = -1;                  // -1 where it applies, 0 off, 1 required
  bool sweep_flow = false;            // in use
  int flow_waves = 0;
  int flow_barrier = 1 << 12;         // polls for an octant's waves to arrive
  int flow_early_req = -1;            // early upwind loads: -1 by size, 0 off, 1 on
  int flow_strips = 1;                // bands of y an octant sweeps one after another
  bool flow_early = false;
  TaskRec* flow_rec = nullptr;        // octant-major tasks (flow_setup)
  int64_t flow_off[9] = {};
  uint32_t* flow_flags = nullptr;     // [octant][slot]
  uint32_t* flow_epoch = nullptr;     // the launch's state (k_flow_epoch)
  uint32_t* flow_err = nullptr;
  uint32_t* flow_diag = nullptr;      // per block: octant and tasks done (the failure message)
  int multi_rb = 0;            // right-hand sides per wave of the several-Kn sweep (0: 1,
                               // measured fastest: 1, 2, 4 took 14.1, 14.8, 16.1 s on 16 Kn)
  int oct_groups = 0;          // octant groups of the sweep (1, 2, 4, 8; 0: 4 on a context
                               // with its own stream, 1 on a caller's stream)
  bool oct_serial = false;     // groups one after the other instead of concurrent chains
  hipStream_t side[7] = {};    // streams of octant groups 1.. (concurrent groups)
  hipEvent_t fork_ev = nullptr, join_ev[7] = {};
  bool planes_fused = true;    // inlet/outlet and symmetry planes in one launch (k_planes)
  int32_t* surface = nullptr;  // Geometry::surface
  double* inv_den = nullptr;   // [first-order flags][velocity]
  SweepCoef* coef = nullptr;   // [first-order flags]
  int32_t* hp_task = nullptr;
  SweepWin* hp_win = nullptr;  // storage window per (hyperplane, octant)
  // tiled sweep (k_tile): one launch per tile diagonal instead of one per
  // hyperplane, for grids of at most 64 velocities per octant; selected with
  // PIKS_SWEEP_TILES=1 (measured slower than the hyperplane sweep on the
  // 200^3 sphere pack for every tile shape tried)
  bool sweep_tiles = false;
  int tile_tx = 2, tile_ty = 2, tile_tz = 2;
  uint32_t* tl_rec = nullptr;
  void* view_args[2] = {nullptr, nullptr};  // per precision, device copies for the sweep
  int32_t* pos[4] = {nullptr, nullptr, nullptr, nullptr};
  // FP64 defect correction (PIKS_PREC_FP32_REFINED contexts).  The solve
  // runs in blocks of dc_len iterations: one FP64 iteration of the state S
  // that also yields the residual r = T(S) - S of the iteration map T (the
  // swept slots' r comes from the sweep itself, the other slots' from
  // ns_old, the copy taken before the iteration), then dc_len - 1 FP32
  // iterations of the correction e <- A(e + r) with the linear part A of T
  // (pressure drop 0), kept in s32 with the moments in mom_e.  In exact
  // arithmetic S + e + r after j corrections is the plain iteration's state
  // j + 1 iterations after the block began (T^k(S) = S + sum_{i<k} A^i r), so
  // FP32 rounds only the correction; the state the host sees is S + s32
  // (dc_pending) or S, folded into S by the next block or by ctx_dc_flush.
  // dc_len follows the convergence check (ctx_dc_policy): 1 (plain FP64)
  // while the per-iteration change is large, longer blocks as it falls.
  State<float> r32;
  double* ns_old = nullptr;          // [rhs][non-swept slot][velocity]
  double *fw_old = nullptr, *bufe_old = nullptr, *bufw_old = nullptr;
  Mom *mom_e = nullptr, *mom_old = nullptr, *mom_r = nullptr;
  uint32_t* ns_list = nullptr;       // the non-swept slots of f (build_nonswept)
  uint32_t* ns_act = nullptr;        // [rhs][n_ns]: the block's slots with a nonzero residual
  uint32_t* ns_act_n = nullptr;      // [rhs][wave of k_dc_slots]: the wave's count of them
  bool dc_active = true;             // correction iterations walk ns_act instead of ns_list
  std::vector<uint32_t> ns_host;
  int64_t n_ns = 0;
  int dc_len = 1, dc_pos = 0, dc_fixed = 0;
  bool dc_pending = false;
  int64_t dc_inner = 0;              // FP32 correction iterations run
  // one iteration of each kind (plain FP64, the defect-correction block's
  // FP64 iteration without and with folding, the first and later correction
  // iterations) captured as a HIP graph (used by ctx_enqueue when use_graph
  // is set)
  bool use_graph = false;
  hipGraphExec_t graph[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};  // per iteration kind
  // timing: four events per step of the last timed call
  std::vector<hipEvent_t> ev;
  double last_ms[4] = {0, 0, 0, 0};
  std::vector<void*> allocs;
};

// Records the calling thread's last error message and returns code.
int set_error(int code, const std::string& msg);

// Host phase of set-up: parameter checks, velocity grid, image preprocessing.
int ctx_create_host(const piks_params* p, const uint8_t* raw, int64_t nbytes, Ctx** out);
// Device memory ctx_create_device will allocate (bytes).
size_t ctx_device_bytes(const Ctx& c);
// Device phase: allocate, clear and upload.  stream == nullptr: the context
// creates its own stream and uses hipMalloc; otherwise it works on `stream`
// with stream-ordered allocations.  Returns after the uploads have completed
// (it synchronises the stream it works on).
int ctx_create_device(Ctx& c, hipStream_t stream);
// Enqueue nsteps iterations (no synchronisation, no timing).
int ctx_enqueue(Ctx& c, int nsteps);
// Enqueue the copy of Ux on the plane x = 2 (Ny*Nz doubles per right-hand
// side, j fastest) into host memory (pinned, so the copy is asynchronous).
int ctx_enqueue_plane(Ctx& c, double* host_plane);
// Mass flow through x = 2 from one right-hand side's plane, in the
// reference's summation order (massLocal of src/output.F90:80-142).
double ctx_mass(const Ctx& c, const double* plane);
// chkConverge's update of the stored mass of right-hand side r
// (src/output.F90:151-155): out = {mass, permeability, error}.
void ctx_check(Ctx& c, int r, double m, double out[3]);
// Set the right-hand sides before ctx_create_device: nrhs Knudsen numbers.
int ctx_set_rhs(Ctx& c, int nrhs, const double* kn);
// Switch between plain FP64 iterations and defect correction (refined
// contexts only); a pending correction is folded into the FP64 state first.
int ctx_set_phase(Ctx& c, int phase);
// Fold a pending correction into the FP64 state (enqueued on the stream).
int ctx_dc_flush(Ctx& c);
// Block length of the defect correction after a check whose relative change
// per iteration was err (the largest over the right-hand sides).
void ctx_dc_policy(Ctx& c, double err);
// Free everything (stream-ordered frees on a caller's stream do not block).
void ctx_destroy(Ctx* c);

}  // namespace piks
