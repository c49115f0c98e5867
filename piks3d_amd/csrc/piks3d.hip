// piks3d.hip -- MI355X (gfx950) DVM iteration for linearized BGK flow through a
// voxel image, behind the C ABI in include/piks3d.h.
//
// One source iteration of the reference (src/solver.F90:111-1940) becomes:
//   sweep      8 octants at once; one launch per oriented hyperplane
//              i+j+k; a 64-lane wave updates all velocities of one (cell,
//              octant) (f loads coalesced along the velocity index); partial
//              moments of each (cell, octant) reduced in registers
//   commit     the reference's in-sweep wall switches, applied after the sweep
//              (readers were redirected to the switched sources during it)
//   halo       periodic x exchange with the reference's one-iteration lag
//   wall BC    one wave per wall: extrapolate, reduce the wall flux over the
//              velocities, re-emit (diffuse + specular)
//   in/out     pressure-drop inlet/outlet, then the four symmetry planes
//   moments    sum of the sweep partials; recomputed from f on boundary planes
// Device layout: f[octant][cell][velocity] over the full ghosted grid, cell =
// (i+1) + (j+1)Nxt + (k+1)Nxt*Nyt, velocity fastest.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/piks3d.h"
#include "context.h"
#include "gauss_hermite_tables.h"
#include "geometry.h"
#include "wallprog.h"

using namespace piks;

namespace piks {

thread_local std::string g_err;

int set_error(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

int flow_check(Ctx& c);  // the single-launch sweep's failure flag (defined below)

}  // namespace piks

namespace {

int fail(int code, const std::string& msg) { return set_error(code, msg); }

#define HIP_TRY(expr)                                                                  \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess)                                                              \
      return fail(PIKS_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));      \
  } while (0)

constexpr int kWave = 64;
constexpr int kMaxTimedSteps = 4096;  // steps of one piks_iterate call timed by events

// ------------------------------------------------------------------ helpers

// Lane exchanges of 64-bit values in registers (no LDS round trip):
// swap_halves(x, y) leaves x = [x lanes 0-31, y lanes 0-31] and y = [x lanes
// 32-63, y lanes 32-63] (gfx950 v_permlane32_swap); swap_rows(x, y) leaves
// x = [x row 0, y row 0, x row 2, y row 2] and y = [x row 1, y row 1, x row 3,
// y row 3] for 16-lane rows (v_permlane16_swap); dpp<C>(v) reads v from the
// lane that the DPP control C names inside the lane's row.
__device__ __forceinline__ void swap_halves(double& x, double& y) {
  const uint64_t ux = __builtin_bit_cast(uint64_t, x), uy = __builtin_bit_cast(uint64_t, y);
  const auto lo = __builtin_amdgcn_permlane32_swap((uint32_t)ux, (uint32_t)uy, false, false);
  const auto hi = __builtin_amdgcn_permlane32_swap((uint32_t)(ux >> 32), (uint32_t)(uy >> 32), false, false);
  x = __builtin_bit_cast(double, ((uint64_t)hi[0] << 32) | lo[0]);
  y = __builtin_bit_cast(double, ((uint64_t)hi[1] << 32) | lo[1]);
}
__device__ __forceinline__ void swap_rows(double& x, double& y) {
  const uint64_t ux = __builtin_bit_cast(uint64_t, x), uy = __builtin_bit_cast(uint64_t, y);
  const auto lo = __builtin_amdgcn_permlane16_swap((uint32_t)ux, (uint32_t)uy, false, false);
  const auto hi = __builtin_amdgcn_permlane16_swap((uint32_t)(ux >> 32), (uint32_t)(uy >> 32), false, false);
  x = __builtin_bit_cast(double, ((uint64_t)hi[0] << 32) | lo[0]);
  y = __builtin_bit_cast(double, ((uint64_t)hi[1] << 32) | lo[1]);
}
template <int kCtrl>
__device__ __forceinline__ double dpp(double v) {
  const uint64_t u = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_mov_dpp((int)(uint32_t)u, kCtrl, 0xf, 0xf, false);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_mov_dpp((int)(uint32_t)(u >> 32), kCtrl, 0xf, 0xf, false);
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}
constexpr int kDppRowRor8 = 0x128, kDppRowRor4 = 0x124;   // rotate the row by 8, 4 lanes
constexpr int kDppXor2 = 0x4e, kDppXor1 = 0xb1;           // quad_perm [2,3,0,1], [1,0,3,2]

// Nibble o, bits 1-3: the octant's x, y, z signs are negative (bit 0, the
// density, never is).
constexpr uint32_t oct_neg_bits() {
  uint32_t b = 0;
  for (int o = 0; o < 8; ++o)
    b |= (uint32_t)((osx(o) < 0 ? 2 : 0) | (osy(o) < 0 ? 4 : 0) | (osz(o) < 0 ? 8 : 0)) << (4 * o);
  return b;
}
constexpr uint32_t kOctNeg = oct_neg_bits();
static_assert(kOctNeg == 0xCEA84620u, "octant sign table");

// Inside a 16-lane row whose values repeat with period 16 / 2^k, the row
// rotations by 8 and 4 read the same value as the partner lane ^ 8 and ^ 4, so
// the four steps below form the butterfly over lane ^ 8, ^ 4, ^ 2, ^ 1.
__device__ __forceinline__ double row_sum(double z) {
  z += dpp<kDppRowRor8>(z);
  z += dpp<kDppRowRor4>(z);
  z += dpp<kDppXor2>(z);
  z += dpp<kDppXor1>(z);
  return z;
}

// Sum over the wave, in every lane: the butterfly over lane ^ 32, ^ 16, ^ 8,
// ^ 4, ^ 2, ^ 1 (each step adds the partner's value to the lane's own).
__device__ __forceinline__ double wave_sum(double v) {
  double x = v, y = v;
  swap_halves(x, y);
  double z = x + y;
  x = z; y = z;
  swap_rows(x, y);
  return row_sum(x + y);
}

__device__ __forceinline__ int64_t uniform64(int64_t v) {
  int lo = __builtin_amdgcn_readfirstlane((int)(v & 0xffffffff));
  int hi = __builtin_amdgcn_readfirstlane((int)(v >> 32));
  return ((int64_t)hi << 32) | (uint32_t)lo;
}

// Read-only data of a kernel loaded through the scalar cache: the constant
// address space tells the compiler that no store of the kernel can clobber it.
template <typename T>
__device__ __forceinline__ T ld_const(const T* p) {
#if defined(__HIP_DEVICE_COMPILE__)
  return *(const __attribute__((address_space(4))) T*)p;
#else
  return *p;  // host pass: never called
#endif
}

// element offset of (octant o, cell, velocity 0) in f
__device__ __forceinline__ size_t fat(const Dims& d, int o, int64_t cell) {
  return (size_t)o * d.ostride + (size_t)d.pos[ofam(o)][cell] * d.Nc8;
}

// ------------------------------------------------------------------ sweep

template <typename T>
struct ViewArgs;

template <typename T>
struct SweepArgs {
  T* f;
  const Mom* mom;
  double* part;            // [cell][octant][4] partial moments; null = skip
  const TaskRec* tasks;    // this hyperplane's task records
  const SweepWin* win;     // this hyperplane's storage window per octant
  const double* cx;
  const double* cy;
  const double* cz;
  const double* w;
  const double* mu;        // per right-hand side: sqrt(pi/2) / Kn
  double mu0;              // mu of right-hand side 0 (the only one of a single solve)
  const double* inv_den;   // [first-order flags][velocity] 1 / denominator, right-hand side 0
  const SweepCoef* coef;   // [first-order flags]
  double c0, c1, c2, d1;   // 1.5/ds, 2/ds, -0.5/ds, 1/ds
  const ViewArgs<T>* view; // what the rare switched-wall reads need
  int64_t rs_f, rs_mom;    // strides between right-hand sides (elements)
  int32_t nrhs;
  int32_t ntask;
  int32_t Nc8;
  uint32_t xcd_per;        // blocks per XCD (grid / 8) when renumbering, else 0
  // defect correction (kernels with kDC != 0; same layout and windows as f):
  // kDC = 1, the FP64 iteration of a block: the previous value of a swept
  // slot is f + dc_e when dc_fold (the block before left a correction), and
  // dc_r receives new - previous; kDC = 2, an FP32 correction iteration (f
  // is the correction e): the previous value is e + dc_r, or dc_r alone in
  // the block's first correction iteration (dc_first)
  float* dc_r;
  const float* dc_e;
  int32_t dc_fold, dc_first;
};

// Everything a read of a switched wall slot needs (device memory).
template <typename T>
struct ViewArgs {
  const T* f;
  const T* fw;             // [corner][octant][velocity]
  const uint8_t* lab;
  const int32_t* wid;
  const uint64_t* views;
  const int32_t* wall_corner;
  Dims d;
};


// Value a cell reading neighbour n along `axis` (1 = y, 2 = z) sees in octant
// o, when the reference would have switched that wall slot by then; null if
// the plain f slot applies.
template <typename T>
__device__ __forceinline__ const T* switched_source(const ViewArgs<T>& a, const T* f, const T* fw, int64_t n,
                                                    int o, int axis) {
  if (a.lab[n] < 20) return nullptr;
  const int w = a.wid[n];
  const int src = (int)((a.views[w] >> (8 * o + (axis == 2 ? 4 : 0))) & 15u);
  if (src == 0) return nullptr;
  if (src <= 8) return f + fat(a.d, src - 1, n);
  return fw + ((size_t)a.wall_corner[w] * 8 + o) * a.d.Nc8;
}

// Sum of four per-lane values over the wave: lane 0 ends with the sum of a,
// lane 16 of b, lane 32 of c, lane 48 of d.  The same butterfly as wave_sum
// (lane ^ 32, then ^ 16, ^ 8, ^ 4, ^ 2, ^ 1), with the four sums sharing its
// first two steps: after swap_halves the lower half adds a's halves and the
// upper half c's (and b's, d's in the second pair), after swap_rows the rows
// hold a, b, c, d partials.
__device__ __forceinline__ double wave_sum4(double a, double b, double c, double d) {
  swap_halves(a, c);
  swap_halves(b, d);
  double x = a + c, y = b + d;
  swap_rows(x, y);
  return row_sum(x + y);
}

// Workgroups are dispatched round-robin over the eight XCDs (block b runs on
// XCD b % 8).  Renumber them so that XCD x runs the x-th contiguous eighth of
// the launch's tasks: consecutive tasks of one octant share upwind
// neighbours, which then stay in that XCD's L2.  Launch grids are multiples
// of 8 blocks; per = grid / 8 comes with the arguments (0: no renumbering).
__device__ __forceinline__ uint32_t xcd_block(uint32_t per) {
  return per ? (blockIdx.x % 8) * per + blockIdx.x / 8 : blockIdx.x;
}

// Index of the calling wave inside its workgroup, marked wave-uniform: every
// per-item load and branch that follows then runs on the scalar unit.
__device__ __forceinline__ int wave_in_block() {
  return __builtin_amdgcn_readfirstlane((int)(threadIdx.x / kWave));
}

// Raw (untyped) buffer resources, word 3 for the gfx9 family: loads outside
// the resource's byte range return 0 and stores outside it are dropped.
constexpr int kBufferWord3 = 0x00020000;

template <typename T>
__device__ __forceinline__ T buf_load(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff);
template <>
__device__ __forceinline__ double buf_load<double>(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}
template <>
__device__ __forceinline__ float buf_load<float>(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}
typedef unsigned int u32x2 __attribute__((__vector_size__(2 * sizeof(unsigned int))));
__device__ __forceinline__ void buf_store(double v, __amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, voff, soff, 0);
}
__device__ __forceinline__ void buf_store(float v, __amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, voff, soff, 0);
}

// Values seen by a cell whose y or z upwind neighbours are switched wall
// slots (few tasks take this path).  view_fix is the out-of-line call;
// view_fix_body the same code for callers that inline it (the FP64
// single-solve sweep: 4.6 % faster inlined, the FP32 one is not).
template <typename T>
struct ViewFix {
  T v[4];                  // y1 y2 z1 z2
  uint32_t mask;           // which of them replace the plain slots
};

template <typename T>
__device__ __forceinline__ ViewFix<T> view_fix_body(const ViewArgs<T>* va, uint32_t cell, int o, int ll,
                                                    int rhs) {
  const ViewArgs<T>& a = *va;
  const T* f = a.f + rhs * a.d.rs_f;
  const T* fw = a.fw + rhs * a.d.rs_fw;
  const int64_t dy = (int64_t)osy(o) * a.d.Nxt, dz = (int64_t)osz(o) * a.d.Nxyt;
  const int64_t nbr[4] = {(int64_t)cell - dy, (int64_t)cell - 2 * dy, (int64_t)cell - dz, (int64_t)cell - 2 * dz};
  ViewFix<T> r{};
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const T* src = switched_source(a, f, fw, nbr[q], o, q < 2 ? 1 : 2);
    if (src) {
      r.v[q] = src[ll];
      r.mask |= 1u << q;
    }
  }
  return r;
}

template <typename T>
__device__ __noinline__ ViewFix<T> view_fix(const ViewArgs<T>* va, uint32_t cell, int o, int ll, int rhs) {
  return view_fix_body(va, cell, o, ll, rhs);
}

// Previous value of a swept slot (byte offset `self` in f's window) as the
// update reads it; with defect correction (kDC, see SweepArgs) the stored
// value plus the correction terms.
template <typename T, int kDC>
__device__ __forceinline__ double dc_previous(const SweepArgs<T>& a, __amdgpu_buffer_rsrc_t rs,
                                              __amdgpu_buffer_rsrc_t rsr, __amdgpu_buffer_rsrc_t rse,
                                              uint32_t voff, uint32_t self) {
  const double v = buf_load<T>(rs, voff, self);
  if constexpr (kDC == 1) {
    // FP64 iteration of a block: fold the previous block's correction
    const uint32_t s4 = self * 4u / (uint32_t)sizeof(T), v4 = voff * 4u / (uint32_t)sizeof(T);
    return a.dc_fold ? v + (double)buf_load<float>(rse, v4, s4) : v;
  } else if constexpr (kDC == 2) {
    // correction iteration: e + r (e is f here), r alone the first time
    const double r = buf_load<float>(rsr, voff, self);
    return a.dc_first ? r : v + r;
  }
  return v;
}

// One hyperplane of the eight octant sweeps.  Every task (cell, octant) has
// its upwind neighbours on the two previous hyperplanes (earlier launches), so
// no wave waits on another.  A wave updates one task, all its velocities (64
// per pass).  The task record holds the element offsets of the cell and of its
// six upwind neighbours inside the octant's storage window for this
// hyperplane, so every f access is a buffer access with a scalar offset and
// the lane's velocity as vector offset; the zero slot is encoded as the
// window's end, where the buffer returns 0.
// Each update is src/solver.F90:206-214 (and analogues) with the same
// operations in the same order (contraction off): signs enter as exact
// negations, and a first-order axis contributes (c*0)*f like the reference.
template <typename T, int kThreads, bool kRecip, int RB, int kDC = 0>
__global__ __launch_bounds__(kThreads) void k_sweep(SweepArgs<T> a) {
  // the update runs in FP64 in both phases: FP32 arithmetic leaves a larger
  // error along the iteration's slowest mode than FP32 storage alone
  using C = double;
  // the wave index is uniform over the wave: keep it (and everything derived
  // from it: the record, cell, octant, the uniform parts of addresses) in SGPRs
  const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  const uint32_t t = xcd_block(a.xcd_per) * (kThreads / kWave) + wave;
  if (t >= (uint32_t)a.ntask) return;
  const int lane = threadIdx.x & (kWave - 1);
  const TaskRec r = ld_const(a.tasks + t);
  const int o = (int)((uint32_t)r.w[0] >> 28);
  const uint32_t self = ((uint32_t)r.w[0] & 0x0fffffffu) * (uint32_t)sizeof(T);
  const uint32_t cell = (uint32_t)r.w[7] & 0x0fffffffu;
  const uint32_t flg = (uint32_t)r.w[7] >> 28;
  const SweepWin win = ld_const(a.win + o);
  // this wave's right-hand sides: RB of them from r0 (the record is shared)
  const int r0 = blockIdx.y * RB;
  const int nr = a.nrhs - r0 < RB ? a.nrhs - r0 : RB;
  Mom m[RB];
  C mu[RB];
  __amdgpu_buffer_rsrc_t rs[RB], rsr[RB], rse[RB];
#pragma unroll
  for (int k = 0; k < RB; ++k) {
    const int rr = k < nr ? r0 + k : r0;
    m[k] = ld_const(a.mom + rr * a.rs_mom + cell);
    mu[k] = a.nrhs == 1 ? a.mu0 : ld_const(a.mu + rr);
    rs[k] = __builtin_amdgcn_make_buffer_rsrc(a.f + rr * a.rs_f + win.off, 0,
                                              (int)(win.len * (uint32_t)sizeof(T)), kBufferWord3);
    if constexpr (kDC != 0) {
      rsr[k] = __builtin_amdgcn_make_buffer_rsrc(a.dc_r + rr * a.rs_f + win.off, 0, (int)(win.len * 4u), kBufferWord3);
      rse[k] = __builtin_amdgcn_make_buffer_rsrc((float*)a.dc_e + rr * a.rs_f + win.off, 0, (int)(win.len * 4u),
                                                 kBufferWord3);
    } else {
      rsr[k] = rse[k] = rs[k];  // not read without defect correction
    }
  }
  const int sx = osx(o), sy = osy(o), sz = osz(o);
  const int Nc8 = a.Nc8;
  const bool fx = flg & 1u, fy = flg & 2u, fz = flg & 4u;
  for (int base = 0; base < Nc8; base += kWave) {
    const int l = base + lane;
    const bool lact = l < Nc8;
    const int ll = lact ? l : 0;
    const uint32_t voff = (uint32_t)ll * (uint32_t)sizeof(T);
    T nb[RB][6];
    C fp[RB];
#pragma unroll
    for (int k = 0; k < RB; ++k) {
      fp[k] = dc_previous<T, kDC>(a, rs[k], rsr[k], rse[k], voff, self);
#pragma unroll
      for (int q = 0; q < 6; ++q) nb[k][q] = buf_load<T>(rs[k], voff, (uint32_t)r.w[1 + q] * (uint32_t)sizeof(T));
    }
    const C cxl = a.cx[ll], cyl = a.cy[ll], czl = a.cz[ll], wl = a.w[ll];
    const C d1 = a.d1, c0 = a.c0, c1 = a.c1, c2 = a.c2;
#pragma unroll
    for (int k = 0; k < RB; ++k) {
      if (k >= nr) break;
      if (flg & 8u) {  // reads a wall slot the reference switched
        const ViewFix<T> vf = view_fix(a.view, cell, o, ll, r0 + k);
#pragma unroll
        for (int q = 0; q < 4; ++q)
          if (vf.mask & (1u << q)) nb[k][2 + q] = vf.v[q];
      }
      C tt = sx > 0 ? cxl * m[k].x : -(cxl * m[k].x);
      tt = sy > 0 ? tt + cyl * m[k].y : tt - cyl * m[k].y;
      tt = sz > 0 ? tt + czl * m[k].z : tt - czl * m[k].z;
      const C feq = wl * (m[k].r + C(1) * tt);
      C num = mu[k] * (feq - C(0.5) * fp[k]);
      num = num + cxl * (fx ? d1 : c1) * nb[k][0];
      num = num + cxl * (fx ? C(0) : c2) * nb[k][1];
      num = num + cyl * (fy ? d1 : c1) * nb[k][2];
      num = num + cyl * (fy ? C(0) : c2) * nb[k][3];
      num = num + czl * (fz ? d1 : c1) * nb[k][4];
      num = num + czl * (fz ? C(0) : c2) * nb[k][5];
      C fnew;
      if (kRecip) {  // 1 / denominator from the table (single right-hand side)
        fnew = num * a.inv_den[(flg & 7u) * (uint32_t)Nc8 + (uint32_t)ll];
      } else {
        C den = C(0.5) * mu[k] + cxl * (fx ? d1 : c0);
        den = den + cyl * (fy ? d1 : c0);
        den = den + czl * (fz ? d1 : c0);
        fnew = num / den;
      }
      if (lact) {
        buf_store((T)fnew, rs[k], voff, self);
        if constexpr (kDC == 1) buf_store((float)(fnew - fp[k]), rsr[k], (uint32_t)ll * 4u, self * 4u / (uint32_t)sizeof(T));
      }
      if (a.part) {
        // moments are summed in FP64 in both phases
        const double fv = lact ? (double)(T)fnew : 0.0;
        const double z = wave_sum4(fv, cxl * fv, cyl * fv, czl * fv);
        if ((lane & 15) == 0) {
          const int q = lane >> 4;
          const double sgn = q == 1 ? sx : q == 2 ? sy : q == 3 ? sz : 1;
          double* pp = a.part + (size_t)(r0 + k) * a.rs_mom * 32 + ((size_t)cell * 8 + o) * 4 + q;
          // velocity passes after the first add to the first pass's sum
          *pp = base == 0 ? sgn * z : *pp + sgn * z;
        }
      }
    }
  }
}

// The single-solve sweep with the least scalar work per task: the
// flag-dependent coefficients read from a table, the octant's signs applied
// to the moments (an exact negation) instead of branching.  A task's wave
// spends most of its life waiting on a chain of dependent loads (its record,
// then the cell's moments and the upwind values, then the store).  With
// NT > 1 a wave takes NT consecutive tasks of the hyperplane and issues every
// load of all of them before the first update (the tasks of one hyperplane
// never read each other's slots: upwind neighbours lie on the two previous
// hyperplanes, so the values are unchanged); on the 200^3 sphere pack NT = 2,
// 3, 4 took 22.3, 24.3, 25.9 ms per sweep against 20.2 for NT = 1, the
// default: more requests in flight per wave do not shorten a sweep whose
// upwind values are re-fetched after every launch.  k_sweep1 serves grids of
// at most 64 velocities per octant in one pass of the wave; k_sweep1_passes
// (NT = 1) loops over passes of 64 velocities for larger grids.
template <typename T, bool kRecip, bool kOnePass, int kDC = 0, int NT = 1, bool kMulti = false>
__device__ __forceinline__ void sweep1_task(const SweepArgs<T>& a) {
  static_assert(kOnePass || NT == 1, "velocity passes take one task per wave");
  static_assert(!kMulti || (!kRecip && NT == 1), "several right-hand sides: one task per wave, no table");
  // the right-hand side of a several-Kn context (blockIdx.y; its own state,
  // moments, partials and mu), right-hand side 0 of a single solve
  const int rhs = kMulti ? (int)blockIdx.y : 0;
  T* const fr = kMulti ? a.f + (size_t)rhs * a.rs_f : a.f;
  const Mom* const momr = kMulti ? a.mom + (size_t)rhs * a.rs_mom : a.mom;
  double* const partr = kMulti && a.part ? a.part + (size_t)rhs * a.rs_mom * 32 : a.part;
  float* const dcr = kMulti ? a.dc_r + (size_t)rhs * a.rs_f : a.dc_r;
  const float* const dce = kMulti ? a.dc_e + (size_t)rhs * a.rs_f : a.dc_e;
  const double mu0 = kMulti ? ld_const(a.mu + rhs) : a.mu0;
  // the padding tasks (t >= ntask) run the last task without storing
  // anything, so that no load waits behind the bounds test
  const uint32_t t0 = xcd_block(a.xcd_per) * (uint32_t)NT;
  const int lane = threadIdx.x;
  bool live[NT];
  uint32_t o[NT], self[NT], cell[NT], flg[NT];
  Mom m[NT];
  SweepCoef cf[NT];
  __amdgpu_buffer_rsrc_t rs[NT], rsr[NT], rse[NT];
  int32_t nbo[NT][6];
#pragma unroll
  for (int j = 0; j < NT; ++j) {
    const uint32_t t = t0 + (uint32_t)j;
    live[j] = t < (uint32_t)a.ntask;
    const TaskRec r = ld_const(a.tasks + (live[j] ? t : (uint32_t)a.ntask - 1u));
    o[j] = (uint32_t)r.w[0] >> 28;
    self[j] = ((uint32_t)r.w[0] & 0x0fffffffu) * (uint32_t)sizeof(T);
    cell[j] = (uint32_t)r.w[7] & 0x0fffffffu;
    flg[j] = (uint32_t)r.w[7] >> 28;
#pragma unroll
    for (int q = 0; q < 6; ++q) nbo[j][q] = r.w[1 + q];
    const SweepWin win = ld_const(a.win + o[j]);
    m[j] = ld_const(momr + cell[j]);
    cf[j] = ld_const(a.coef + (flg[j] & 7u));
    rs[j] = __builtin_amdgcn_make_buffer_rsrc(fr + win.off, 0, (int)(win.len * (uint32_t)sizeof(T)), kBufferWord3);
    rsr[j] = rse[j] = rs[j];
    if constexpr (kDC != 0) {
      rsr[j] = __builtin_amdgcn_make_buffer_rsrc(dcr + win.off, 0, (int)(win.len * 4u), kBufferWord3);
      rse[j] = __builtin_amdgcn_make_buffer_rsrc((float*)dce + win.off, 0, (int)(win.len * 4u), kBufferWord3);
    }
  }
  const int Nc8 = a.Nc8;
  // lane l holds velocity base + l of the octant; one pass when Nc8 <= 64
  for (int base = 0; base < (kOnePass ? 1 : Nc8); base += kWave) {
    const int l = base + lane;
    const bool lact = l < Nc8;
    const int ll = lact ? l : 0;
    const uint32_t voff = (uint32_t)ll * (uint32_t)sizeof(T);
    double fp[NT], nb[NT][6];
#pragma unroll
    for (int j = 0; j < NT; ++j) {
      fp[j] = dc_previous<T, kDC>(a, rs[j], rsr[j], rse[j], voff, self[j]);
#pragma unroll
      for (int q = 0; q < 6; ++q) nb[j][q] = buf_load<T>(rs[j], voff, (uint32_t)nbo[j][q] * (uint32_t)sizeof(T));
    }
    // FP32: the velocity table is read before the switched-wall call (the
    // compiler moves no load across it), so that these reads overlap the f
    // reads (4 % faster); FP64 reads it after the call (1 % faster there)
    double cxl, cyl, czl, wl;
    if constexpr (sizeof(T) == 4) { cxl = a.cx[ll]; cyl = a.cy[ll]; czl = a.cz[ll]; wl = a.w[ll]; }
    double kinv[NT];
#pragma unroll
    for (int j = 0; j < NT; ++j) kinv[j] = kRecip ? a.inv_den[(flg[j] & 7u) * (uint32_t)Nc8 + (uint32_t)ll] : 0.0;
#pragma unroll
    for (int j = 0; j < NT; ++j) {
      if (flg[j] & 8u) {  // reads a wall slot the reference switched
        const ViewFix<T> vf = sizeof(T) == 8 && NT == 1 ? view_fix_body(a.view, cell[j], (int)o[j], ll, rhs)
                                                        : view_fix(a.view, cell[j], (int)o[j], ll, rhs);
#pragma unroll
        for (int q = 0; q < 4; ++q)
          if (vf.mask & (1u << q)) nb[j][2 + q] = vf.v[q];
      }
    }
    if constexpr (sizeof(T) == 8) { cxl = a.cx[ll]; cyl = a.cy[ll]; czl = a.cz[ll]; wl = a.w[ll]; }
#pragma unroll
    for (int j = 0; j < NT; ++j) {
      // signed moments: cx*(-u) is -(cx*u) exactly, and t - y is t + (-y)
      const double mxj = osx((int)o[j]) > 0 ? m[j].x : -m[j].x;
      const double myj = osy((int)o[j]) > 0 ? m[j].y : -m[j].y;
      const double mzj = osz((int)o[j]) > 0 ? m[j].z : -m[j].z;
      double tt = cxl * mxj;
      tt = tt + cyl * myj;
      tt = tt + czl * mzj;
      const double feq = wl * (m[j].r + 1.0 * tt);
      double num = mu0 * (feq - 0.5 * fp[j]);
      num = num + cxl * cf[j].x1 * nb[j][0];
      num = num + cxl * cf[j].x2 * nb[j][1];
      num = num + cyl * cf[j].y1 * nb[j][2];
      num = num + cyl * cf[j].y2 * nb[j][3];
      num = num + czl * cf[j].z1 * nb[j][4];
      num = num + czl * cf[j].z2 * nb[j][5];
      double fnew;
      if (kRecip) {
        fnew = num * kinv[j];
      } else {
        double den = 0.5 * mu0 + cxl * cf[j].dx;
        den = den + cyl * cf[j].dy;
        den = den + czl * cf[j].dz;
        fnew = num / den;
      }
      if (lact && live[j]) {
        buf_store((T)fnew, rs[j], voff, self[j]);
        if constexpr (kDC == 1)
          buf_store((float)(fnew - fp[j]), rsr[j], (uint32_t)ll * 4u, self[j] * 4u / (uint32_t)sizeof(T));
      }
      if (partr && live[j]) {
        const double fv = lact ? (double)(T)fnew : 0.0;
        const double z = wave_sum4(fv, cxl * fv, cyl * fv, czl * fv);
        if ((lane & 15) == 0) {
          // the signs of the Ux, Uy, Uz partials (lanes 16, 32, 48) are the
          // octant's: bit q of neg is set when partial q is negated
          const uint32_t q = (uint32_t)lane >> 4;
          const double zz = ((kOctNeg >> (4 * o[j] + q)) & 1u) ? -z : z;
          double* pp = partr + ((size_t)cell[j] * 8 + o[j]) * 4 + q;
          // passes after the first add to the first pass's sum
          *pp = base == 0 ? zz : *pp + zz;
        }
      }
    }
  }
}

template <typename T, bool kRecip, int kDC = 0, int NT = 1>
__global__ __launch_bounds__(64) void k_sweep1(SweepArgs<T> a) { sweep1_task<T, kRecip, true, kDC, NT>(a); }

// the same task for grids of more than 64 velocities per octant: passes of 64
template <typename T, bool kRecip, int kDC = 0>
__global__ __launch_bounds__(64) void k_sweep1_passes(SweepArgs<T> a) { sweep1_task<T, kRecip, false, kDC>(a); }

// the same task for one of several right-hand sides (blockIdx.y), at most 64
// velocities per octant
template <typename T, int kDC = 0>
__global__ __launch_bounds__(64) void k_sweep1_rhs(SweepArgs<T> a) { sweep1_task<T, false, true, kDC, 1, true>(a); }

// ------------------------------------------------------------------ single-launch sweep

// The whole sweep of one source iteration in one launch: no launch gaps
// between hyperplanes (the default for FP64 single solves with 64
// velocities per octant; PIKS_SWEEP_FLOW).  A block sweeps the octant
// numbered like the XCD it runs on (its hardware XCD id), so every producer
// and consumer of an octant's slots share one L2, and takes its arrival
// number w in that octant from a per-XCD counter.  Tasks are in hyperplane
// order (rec, built by the host), so a task's upwind neighbours come earlier.
// When all W waves of an octant have arrived (the set-up probe sizes W to
// what is resident at once), wave w takes tasks w, w + W, w + 2W, ...: the
// earliest unfinished task can always proceed and no wait is circular.  If
// a wave waits `barrier` polls without that happening (kernels of other
// streams holding CUs), the octant switches to claims: every wave, late
// ones included, takes the next task from a per-octant counter, which needs
// no residency (a task waits only on tasks claimed before it by running
// waves) but serialises the claims (8-12x slower when every task is
// claimed).  A finished task sets its slot's flag to the sweep's epoch after
// its stores have reached L2; a task waits for the flags of the nearest
// swept upwind neighbour along each axis (their own waits cover the one
// behind).  Flags are polled with agent-scope loads, which miss the CU's L1
// (an L1 line polled before the flag changed would be served again: a
// workgroup-scope invalidate leaves it).  Values are plain loads served by
// the XCD's L2: within the launch no CU loads a slot's lines before they
// are final (its readers wait for its flag or for the flag of a task that
// waited for it, and its own task's read of the previous value is followed
// by its store, which updates that CU's L1), and the L1 starts the launch
// invalidated.  Every wait gives up after `spin` polls (a safety net: it
// sets err and the wave ends, so the grid always drains); k_flow_done then
// checks that every octant's tasks were done.  The update is sweep1_task's,
// operation for operation; reads of switched wall slots come resolved in
// the records (flow_switched_word).
struct FlowArgs {
  double* f;
  const Mom* mom;
  double* part;             // [cell][octant][4] partial moments; null = skip
  const TaskRec* rec;       // octant-major tasks (flow_setup): word 0 the slot,
                            // words 1..6 the upwind sources (flow_source)
  uint32_t* flags;          // [octant][slot]: epoch of the last update
  const uint32_t* epoch;    // this sweep's epoch (k_flow_epoch)
  uint32_t* err;
  uint32_t* wave_ctr;       // [XCD]: waves started (cleared by k_flow_epoch)
  uint32_t* done;           // [octant]: tasks done
  uint32_t* mode;           // [octant]: 0 undecided, 1 static split, 2 claims
  uint32_t* next;           // [octant]: tasks claimed (claims mode)
  uint32_t waves;           // waves per octant
  uint32_t barrier;         // polls a wave waits for its octant's W waves to arrive
  const double* cx;
  const double* cy;
  const double* cz;
  const double* w;
  const SweepCoef* coef;
  const double* fw;         // corner store [corner][octant][velocity]
  double mu0;
  int64_t off[9];           // first task of each octant in rec
  int64_t ostride;          // elements of f per octant
  int64_t nslot;            // flags per octant (stored slots + the zero slot)
  uint32_t spin;            // polls before a wait gives up
  uint32_t probe;           // nonzero: residency probe only (flow_setup), this many polls
  uint32_t* diag;           // [block]: XCD id, then [grid + block]: tasks done
};

// Next epoch of the flags; the counters of the launch cleared.  st: [0]
// epoch, [1..6] err, [8..15] waves started per XCD, [16..23] tasks done per
// octant, [24] the residency probe's count, [32..39] mode per octant,
// [40..47] tasks claimed per octant.
__global__ void k_flow_epoch(uint32_t* st) {
  if (threadIdx.x == 0) __hip_atomic_fetch_add(st, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (threadIdx.x < 56) st[8 + threadIdx.x] = 0u;
}

// After the sweep: every octant's tasks done (else err, unless already set).
struct FlowCounts {
  uint32_t n[8];
};
__global__ void k_flow_done(uint32_t* st, FlowCounts c) {
  const uint32_t o = threadIdx.x;
  if (o < 8 && st[16 + o] != c.n[o] &&
      __hip_atomic_fetch_add(st + 1, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0u) {
    st[2] = o;
    st[3] = st[16 + o];
    st[4] = 0xffffffffu;
    st[5] = c.n[o];
  }
}

// XCD of every block of a grid of n one-wave blocks (flow_probe)
__global__ void k_flow_probe(uint32_t* out, uint32_t n) {
  if (threadIdx.x == 0 && blockIdx.x < n) out[blockIdx.x] = __builtin_amdgcn_s_getreg((3 << 11) | 20) & 15u;
}

// What a task needs before its upwind values: its record's fields, the
// cell's moments, the slot's previous value and one poll of the flags it
// waits for (one per axis at most; an axis without a wait
// issues no load and reads as ep).
struct FlowTask {
  uint32_t self, cell, flg;
  uint32_t nbs[6];          // the record's source words
  const uint32_t* wt[3];
  uint32_t v[3];
  Mom m;
  double fp;
};

// An agent-scope load (it misses the CU's L1); the caller makes it uniform
// where it compares, so that several polls go out before the first returns.
__device__ __forceinline__ uint32_t flow_poll(const uint32_t* p) {
  return __hip_atomic_load(const_cast<uint32_t*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// One poll of each flag the task waits for.
__device__ __forceinline__ void flow_poll3(FlowTask& t, uint32_t ep) {
#pragma unroll
  for (int ax = 0; ax < 3; ++ax) t.v[ax] = t.wt[ax] ? flow_poll(t.wt[ax]) : ep;
}

// The flags a task waits for, from its record, and one poll of each.
__device__ __forceinline__ void flow_targets(FlowTask& t, const TaskRec& r, const uint32_t* flo, uint32_t ep) {
#pragma unroll
  for (int ax = 0; ax < 3; ++ax) {
    const uint32_t w1 = (uint32_t)r.w[1 + 2 * ax], w2 = (uint32_t)r.w[2 + 2 * ax];
    t.wt[ax] = (w1 >> 28) == 8u   ? flo + (w1 & 0x0fffffffu)
               : (w2 >> 28) == 8u ? flo + (w2 & 0x0fffffffu)
                                  : nullptr;
  }
  flow_poll3(t, ep);
}

// The rest of a task's set-up: record fields, moments, previous value.
__device__ __forceinline__ void flow_fields(FlowTask& t, const TaskRec& r, const Mom* __restrict__ mom,
                                            const double* fo, int lane) {
  t.self = (uint32_t)r.w[0] & 0x0fffffffu;
  t.cell = (uint32_t)r.w[7] & 0x0fffffffu;
  t.flg = (uint32_t)r.w[7] >> 28;
#pragma unroll
  for (int q = 0; q < 6; ++q) t.nbs[q] = (uint32_t)r.w[1 + q];
  t.m = ld_const(mom + t.cell);
  t.fp = fo[(size_t)t.self * 64 + lane];
}

__device__ __forceinline__ void flow_issue(FlowTask& t, const TaskRec& r, const Mom* __restrict__ mom,
                                           const double* fo, const uint32_t* flo, int lane, uint32_t ep) {
  t.self = (uint32_t)r.w[0] & 0x0fffffffu;
  t.cell = (uint32_t)r.w[7] & 0x0fffffffu;
  t.flg = (uint32_t)r.w[7] >> 28;
#pragma unroll
  for (int q = 0; q < 6; ++q) t.nbs[q] = (uint32_t)r.w[1 + q];
#pragma unroll
  for (int ax = 0; ax < 3; ++ax) {
    const uint32_t w1 = t.nbs[2 * ax], w2 = t.nbs[2 * ax + 1];
    t.wt[ax] = (w1 >> 28) == 8u   ? flo + (w1 & 0x0fffffffu)
               : (w2 >> 28) == 8u ? flo + (w2 & 0x0fffffffu)
                                  : nullptr;
  }
  t.m = ld_const(mom + t.cell);
  t.fp = fo[(size_t)t.self * 64 + lane];
  flow_poll3(t, ep);
}

// Whether the polls read the sweep's epoch on every flag the task waits for.
__device__ __forceinline__ bool flow_ready(const FlowTask& t, uint32_t ep) {
  return __builtin_amdgcn_readfirstlane(t.v[0]) == ep && __builtin_amdgcn_readfirstlane(t.v[1]) == ep &&
         __builtin_amdgcn_readfirstlane(t.v[2]) == ep;
}

// The first velocity of an upwind source word of a flow record: bits 0..27
// an index, bits 28..30 x, bit 31 c.  c = 0: slot of octant o ^ x; c = 1,
// x = 0: slot of octant o, whose flag the task waits for; c = 1, x != 0:
// corner store entry (corner * 8 + octant).  The host resolves the reads
// of wall slots the reference switched (switched_source) into the second and
// third forms, so the sweep reads every source through this one path.
__device__ __forceinline__ const double* flow_source(const FlowArgs& a, uint32_t o, uint32_t wq) {
  const uint32_t x = (wq >> 28) & 7u;
  const size_t idx = (size_t)(wq & 0x0fffffffu) * 64;
  return (wq >> 31) && x ? a.fw + idx : a.f + (size_t)(o ^ x) * a.ostride + idx;
}

// Claim the octant's next task (lane 0's value; the caller makes it uniform
// where it uses it, so that the claim overlaps other loads).
__device__ __forceinline__ uint32_t flow_claim(uint32_t* next, int lane) {
  uint32_t v = 0;
  if (lane == 0) v = __hip_atomic_fetch_add(next, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return v;
}

// Set the octant's mode to m unless it is set: the mode that holds.
__device__ __forceinline__ uint32_t flow_decide(uint32_t* mode, uint32_t m, int lane) {
  uint32_t old = 0;
  if (lane == 0)
    __hip_atomic_compare_exchange_strong(mode, &old, m, __ATOMIC_RELAXED, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  old = __builtin_amdgcn_readfirstlane(old);
  return old ? old : m;
}

// Poll until the three flags read ep (false: gave up, or another wave did).
// A wave that gives up records err = 1 and the octant, task, slot, epoch and
// last value it saw in err[1..5] for the host's message.
__device__ __forceinline__ bool flow_wait(FlowTask& t, uint32_t ep, uint32_t* err, uint32_t spin, uint32_t o,
                                          uint32_t k, const uint32_t* flo) {
  for (uint32_t it = 0; it < spin; ++it) {
    const uint32_t v0 = __builtin_amdgcn_readfirstlane(t.v[0]), v1 = __builtin_amdgcn_readfirstlane(t.v[1]),
                   v2 = __builtin_amdgcn_readfirstlane(t.v[2]);
    if (v0 == ep && v1 == ep && v2 == ep) return true;
    if ((it & 63u) == 63u && __builtin_amdgcn_readfirstlane(flow_poll(err)) != 0u) return false;
    // a re-poll waits 16 x 64 cycles: fewer polls in flight leave the memory
    // system to the stores being waited for (40^3 pack, sweep ms for a wait
    // of 1, 2, 4, 8, 16: 0.593, 0.542, 0.490, 0.452, 0.434; 200^3 unchanged)
    __builtin_amdgcn_s_sleep(16);
    flow_poll3(t, ep);
  }
  if (threadIdx.x == 0) {
    const bool b0 = t.v[0] != ep, b1 = t.v[1] != ep;
    err[1] = o;
    err[2] = k;
    const uintptr_t p0 = (uintptr_t)t.wt[0], p1 = (uintptr_t)t.wt[1], p2 = (uintptr_t)t.wt[2];
    err[3] = (uint32_t)(((b0 ? p0 : b1 ? p1 : p2) - (uintptr_t)flo) / sizeof(uint32_t));
    err[4] = ep;
    err[5] = b0 ? t.v[0] : b1 ? t.v[1] : t.v[2];
    __hip_atomic_store(err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  return false;
}

// One wave: its tasks of its XCD's octant (k = w, w + W, ... or claimed),
// pipelined so that a task costs one memory round trip besides its update:
// the next task's flags are polled while this task computes, and after this
// task's stores its moments, previous value and, if those flags were set,
// its upwind values go out before the wait for the stores.  This task's
// flag is set before any poll of the next task is looked at again, so a wave
// never waits for its own task.
// The records, moments and coefficients, which nothing in the launch writes,
// come through restrict arguments, so their loads are scalar loads and the
// kernel runs 6 waves per SIMD (SGPR-bound; its occupancy query says 7).
template <bool kEarly>
__global__ __launch_bounds__(64) void k_sweep_flow(FlowArgs a, const TaskRec* __restrict__ rec,
                                                   const Mom* __restrict__ mom, const SweepCoef* __restrict__ coef) {
  const int lane = threadIdx.x;
  if (a.probe) {  // are all blocks of this grid resident at once?  (st[24] counts them in)
    if (lane == 0) __hip_atomic_fetch_add(a.done + 8, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    uint32_t seen = 0;
    for (uint32_t it = 0; it < a.probe; ++it) {
      seen = __builtin_amdgcn_readfirstlane(flow_poll(a.done + 8));
      if (seen >= gridDim.x) break;
      __builtin_amdgcn_s_sleep(2);
    }
    if (lane == 0) a.diag[blockIdx.x] = seen >= gridDim.x ? 0xffffffffu : seen;
    return;
  }
  const uint32_t o = __builtin_amdgcn_readfirstlane(__builtin_amdgcn_s_getreg((3 << 11) | 20) & 7u);
  uint32_t tk = 0;
  if (lane == 0) tk = __hip_atomic_fetch_add(a.wave_ctr + o, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const uint32_t wv = __builtin_amdgcn_readfirstlane(tk);
  if (lane == 0) a.diag[blockIdx.x] = o;
  // the octant's mode: the static split once its W waves have all arrived
  // (the W-th arrival sets it), claims if a wave waited `barrier` polls first
  uint32_t md = wv + 1 == a.waves ? flow_decide(a.mode + o, 1u, lane) : 0u;
  for (uint32_t it = 0; md == 0u && it < a.barrier; ++it) {
    md = __builtin_amdgcn_readfirstlane(flow_poll(a.mode + o));
    if (md == 0u) __builtin_amdgcn_s_sleep(2);
  }
  if (md == 0u) md = flow_decide(a.mode + o, 2u, lane);
  const bool claims = md == 2u;
  const int64_t base = a.off[o], n = a.off[o + 1] - base;
  // the epoch through the vector path (an agent-scope load): a scalar load
  // could be served by a scalar cache line of the previous kernel
  const uint32_t ep = __builtin_amdgcn_readfirstlane(flow_poll(a.epoch));
  double* const fo = a.f + (size_t)o * a.ostride;
  uint32_t* const flo = a.flags + (size_t)o * a.nslot;
  uint32_t done = 0;
  // the lane's velocity, the same for every task of the wave (64 per octant)
  const double cxl = a.cx[lane], cyl = a.cy[lane], czl = a.cz[lane], wl = a.w[lane];
  const bool px = osx((int)o) > 0, py = osy((int)o) > 0, pz = osz((int)o) > 0;
  uint32_t* const next = a.next + o;
  if constexpr (!kEarly) {
  // static: tasks wv, wv + W, ... (a wave beyond W has none); claims: the
  // next unclaimed task, claimed one task ahead
  int64_t k = claims ? (int64_t)__builtin_amdgcn_readfirstlane(flow_claim(next, lane)) : wv < a.waves ? wv : n;
  if (k < n) {
    FlowTask t;
    flow_issue(t, ld_const(rec + base + k), mom, fo, flo, lane, ep);
    int64_t kn = claims ? (int64_t)__builtin_amdgcn_readfirstlane(flow_claim(next, lane)) : k + a.waves;
    TaskRec rn = ld_const(rec + base + (kn < n ? kn : k));
    while (true) {
      if (!flow_wait(t, ep, a.err, a.spin, o, (uint32_t)k, flo)) break;
      asm volatile("" ::: "memory");  // the upwind loads stay after the polls that cleared them
      double nb[6];
#pragma unroll
      for (int q = 0; q < 6; ++q) nb[q] = flow_source(a, o, t.nbs[q])[lane];
      // the coefficients (eight entries, cache-resident) with the upwind values
      const SweepCoef cf = ld_const(coef + (t.flg & 7u));
      // sweep1_task's update, operation for operation
      const double mx = px ? t.m.x : -t.m.x;
      const double my = py ? t.m.y : -t.m.y;
      const double mz = pz ? t.m.z : -t.m.z;
      double tt = cxl * mx;
      tt = tt + cyl * my;
      tt = tt + czl * mz;
      const double feq = wl * (t.m.r + 1.0 * tt);
      double num = a.mu0 * (feq - 0.5 * t.fp);
      num = num + cxl * cf.x1 * nb[0];
      num = num + cxl * cf.x2 * nb[1];
      num = num + cyl * cf.y1 * nb[2];
      num = num + cyl * cf.y2 * nb[3];
      num = num + czl * cf.z1 * nb[4];
      num = num + czl * cf.z2 * nb[5];
      double den = 0.5 * a.mu0 + cxl * cf.dx;
      den = den + cyl * cf.dy;
      den = den + czl * cf.dz;
      const double fnew = num / den;
      const uint32_t self = t.self, cell = t.cell;
      fo[(size_t)self * 64 + lane] = fnew;
      if (a.part) {
        const double z = wave_sum4(fnew, cxl * fnew, cyl * fnew, czl * fnew);
        if ((lane & 15) == 0) {
          const uint32_t q = (uint32_t)lane >> 4;
          const double zz = ((kOctNeg >> (4 * o + q)) & 1u) ? -z : z;
          a.part[((size_t)cell * 8 + o) * 4 + q] = zz;
        }
      }
      const bool more = kn < n;
      uint32_t knn = 0;
      if (more) {
        flow_issue(t, rn, mom, fo, flo, lane, ep);
        if (claims) knn = flow_claim(next, lane);
      }
      // the slot's values (and partials) reach L2 before its flag
      __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
      if (lane == 0) __hip_atomic_store(flo + self, ep, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      ++done;
      if (!more) break;
      k = kn;
      kn = claims ? (int64_t)__builtin_amdgcn_readfirstlane(knn) : kn + a.waves;
      rn = ld_const(rec + base + (kn < n ? kn : k));
    }
  }
  } else {
  // static: tasks wv, wv + W, ... (a wave beyond W has none); claims: the
  // next unclaimed task
  int64_t k = claims ? (int64_t)__builtin_amdgcn_readfirstlane(flow_claim(next, lane)) : wv < a.waves ? wv : n;
  if (k < n) {
    // task k: its record's fields, moments, previous value and upwind values
    FlowTask t;
    flow_issue(t, ld_const(rec + base + k), mom, fo, flo, lane, ep);
    if (!flow_wait(t, ep, a.err, a.spin, o, (uint32_t)k, flo)) goto out;
    asm volatile("" ::: "memory");  // the upwind loads stay after the polls that cleared them
    double nb[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) nb[q] = flow_source(a, o, t.nbs[q])[lane];
    int64_t kn = claims ? (int64_t)__builtin_amdgcn_readfirstlane(flow_claim(next, lane)) : k + a.waves;
    TaskRec rn = ld_const(rec + base + (kn < n ? kn : k));
    while (true) {
      const bool more = kn < n;
      // the next task's flags, polled while this task computes
      FlowTask tn;
      if (more) flow_targets(tn, rn, flo, ep);
      // the coefficients (eight entries, cache-resident)
      const SweepCoef cf = ld_const(coef + (t.flg & 7u));
      // sweep1_task's update, operation for operation
      const double mx = px ? t.m.x : -t.m.x;
      const double my = py ? t.m.y : -t.m.y;
      const double mz = pz ? t.m.z : -t.m.z;
      double tt = cxl * mx;
      tt = tt + cyl * my;
      tt = tt + czl * mz;
      const double feq = wl * (t.m.r + 1.0 * tt);
      double num = a.mu0 * (feq - 0.5 * t.fp);
      num = num + cxl * cf.x1 * nb[0];
      num = num + cxl * cf.x2 * nb[1];
      num = num + cyl * cf.y1 * nb[2];
      num = num + cyl * cf.y2 * nb[3];
      num = num + czl * cf.z1 * nb[4];
      num = num + czl * cf.z2 * nb[5];
      double den = 0.5 * a.mu0 + cxl * cf.dx;
      den = den + cyl * cf.dy;
      den = den + czl * cf.dz;
      const double fnew = num / den;
      // the next task's polls, read before this task's stores go out (a
      // wait on them after the stores would wait for the stores too)
      const bool early = more && flow_ready(tn, ep);
      const uint32_t self = t.self, cell = t.cell;
      fo[(size_t)self * 64 + lane] = fnew;
      if (a.part) {
        const double z = wave_sum4(fnew, cxl * fnew, cyl * fnew, czl * fnew);
        if ((lane & 15) == 0) {
          const uint32_t q = (uint32_t)lane >> 4;
          const double zz = ((kOctNeg >> (4 * o + q)) & 1u) ? -z : z;
          a.part[((size_t)cell * 8 + o) * 4 + q] = zz;
        }
      }
      // the next task's loads go out with this task's stores: its upwind
      // values too when its flags were already set (it then needs no memory
      // round trip of its own before its update)
      uint32_t knn = 0;
      if (more) {
        flow_fields(tn, rn, mom, fo, lane);
        if (early) {
          asm volatile("" ::: "memory");  // the upwind loads stay after the polls that cleared them
#pragma unroll
          for (int q = 0; q < 6; ++q) nb[q] = flow_source(a, o, tn.nbs[q])[lane];
        }
        if (claims) knn = flow_claim(next, lane);
        else rn = ld_const(rec + base + (kn + a.waves < n ? kn + a.waves : kn));
      }
      // the slot's values (and partials) reach L2 before its flag
      __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
      if (lane == 0) __hip_atomic_store(flo + self, ep, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      ++done;
      if (!more) break;
      t = tn;
      k = kn;
      if (claims) {
        kn = (int64_t)__builtin_amdgcn_readfirstlane(knn);
        rn = ld_const(rec + base + (kn < n ? kn : k));
      } else {
        kn += a.waves;
      }
      if (!early) {  // the next task waits on this one's or others' flags
        if (!flow_wait(t, ep, a.err, a.spin, o, (uint32_t)k, flo)) break;
        asm volatile("" ::: "memory");  // the upwind loads stay after the polls that cleared them
#pragma unroll
        for (int q = 0; q < 6; ++q) nb[q] = flow_source(a, o, t.nbs[q])[lane];
      }
    }
  }
  }
out:
  if (lane == 0) {
    a.diag[gridDim.x + blockIdx.x] = done;
    __hip_atomic_fetch_add(a.done + o, done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// ------------------------------------------------------------------ tiled sweep

template <typename T>
struct TileArgs {
  T* f;                    // right-hand side 0
  const Mom* mom;
  double* part;            // [cell][octant][4] partial moments; null = skip
  const uint32_t* rec;     // this launch's first tile record
  const double* cx;
  const double* cy;
  const double* cz;
  const double* w;
  const double* mu;        // per right-hand side
  double mu0;
  const ViewArgs<T>* view;
  double c0, c1, c2, d1;   // 1.5/ds, 2/ds, -0.5/ds, 1/ds
  int64_t rs_f, rs_mom;    // strides between right-hand sides (elements)
  int64_t ostride;         // elements of f per octant
  int32_t nrhs;
  int32_t ntile;
  int32_t Nc8;
  int32_t tz;              // planes per tile (records per tile: tz + 1)
  uint32_t xcd_per;        // blocks per XCD (grid / 8) when renumbering, else 0
};

__device__ __forceinline__ uint32_t lane_u32(uint32_t v, int lane) {
  return (uint32_t)__builtin_amdgcn_readlane((int)v, lane);
}
__device__ __forceinline__ double lane_f64(uint32_t v, int lane) {
  const uint64_t lo = lane_u32(v, lane), hi = lane_u32(v, lane + 1);
  return __builtin_bit_cast(double, lo | hi << 32);
}

// One tile of one octant per wave: the wave walks the tile's planes in order
// and, inside a plane, its TX * TY cells in raster order, so every upwind
// neighbour of a cell is either an earlier cell of the tile (kept in
// registers: the current plane and the two below it) or a cell of an upwind
// tile of an earlier launch (loaded with the plane: the x and y halo, and the
// two planes below the tile).  Cells the octant does not sweep (walls, the
// inflow layer, ghost planes) keep their value through the sweep and are
// loaded like the halo; cells without storage read as zero (their position is
// the window's end, where the buffer returns 0).  The loads of plane c + 1
// and the record of plane c + 2 are in flight while plane c is computed.
// Every memory instruction of the plane loop is issued unconditionally (the
// stores of slots the octant does not sweep go to an offset outside the
// buffer, which drops them), so that the compiler's count of outstanding
// memory operations stays exact and a wait for plane c's data never waits for
// plane c + 1's loads or for earlier stores.
// Each update is the one of sweep1_task (src/solver.F90:206-214 and
// analogues), the same operations in the same order, and the partial moments
// are reduced the same way, so the results are those of the hyperplane sweep
// bit for bit.  The right-hand side is blockIdx.y.
template <typename T, int TX, int TY>
__global__ __launch_bounds__(64) void k_tile(TileArgs<T> a) {
  constexpr int NS = TX * TY;
  constexpr int NM = (NS + 7) / 8;   // moment vectors: 8 cells (32 B each) per load
  constexpr int W = Geometry::kTileRecDw;
  static_assert(tile_rec_cell(NS, TX, TY, NS) <= W && NS <= 32, "tile record");
  const uint32_t t = xcd_block(a.xcd_per);
  if (t >= (uint32_t)a.ntile) return;
  const int lane = threadIdx.x;
  const int rhs = blockIdx.y;
  const uint32_t* rec = a.rec + (size_t)t * (uint32_t)(a.tz + 1) * W;
  const uint32_t hv = rec[lane];
  const uint32_t h0 = lane_u32(hv, 0);
  const uint32_t o = h0 & 7u;
  const int np = (int)(h0 >> 8);
  const uint32_t lo = lane_u32(hv, 1), len = lane_u32(hv, 2);
  const uint32_t cell_lo = lane_u32(hv, 3), cell_n = lane_u32(hv, 4);
  const int Nc8 = a.Nc8;
  const uint32_t cb = (uint32_t)Nc8 * (uint32_t)sizeof(T);  // bytes per cell
  const uint32_t oob = len * cb;                            // the zero slot; stores there are dropped
  const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
      a.f + rhs * a.rs_f + (size_t)o * a.ostride + (size_t)lo * (uint32_t)Nc8, 0, (int)oob, kBufferWord3);
  // partial moments of the tile's cells: [cell - cell_lo][octant][4] doubles
  const uint32_t poob = cell_n * 256u;
  const __amdgpu_buffer_rsrc_t rp = __builtin_amdgcn_make_buffer_rsrc(
      a.part ? a.part + (size_t)rhs * a.rs_mom * 32 + (size_t)cell_lo * 32 : nullptr, 0, a.part ? (int)poob : 0,
      kBufferWord3);
  const bool lact = lane < Nc8;
  const int ll = lact ? lane : 0;
  const uint32_t voff = (uint32_t)ll * (uint32_t)sizeof(T);
  const uint32_t voff_st = lact ? voff : oob;                 // lanes past the grid store nothing
  const uint32_t voff_p = (lane & 15) == 0 ? (uint32_t)(lane >> 4) * 8u : poob;
  const uint32_t* momw = reinterpret_cast<const uint32_t*>(a.mom + rhs * a.rs_mom);
  const double mu = a.nrhs == 1 ? a.mu0 : a.mu[rhs];
  const double cxl = a.cx[ll], cyl = a.cy[ll], czl = a.cz[ll], wl = a.w[ll];
  const double den0 = 0.5 * mu;

  T p2[NS], p1[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    p2[s] = buf_load<T>(rs, voff, lane_u32(hv, tile_rec_zhalo(NS, 0) + s) * cb);
    p1[s] = buf_load<T>(rs, voff, lane_u32(hv, tile_rec_zhalo(NS, 1) + s) * cb);
  }
  // plane c's values (cn, xn, yn, mn) are issued one plane ahead
  T cn[NS], xn[2 * TY], yn[2 * TX];
  uint32_t mn[NM];
  auto issue = [&](uint32_t r) {
#pragma unroll
    for (int s = 0; s < NS; ++s) cn[s] = buf_load<T>(rs, voff, lane_u32(r, tile_rec_slot(s)) * cb);
#pragma unroll
    for (int q = 0; q < 2 * TY; ++q) xn[q] = buf_load<T>(rs, voff, lane_u32(r, tile_rec_xhalo(NS, q)) * cb);
#pragma unroll
    for (int q = 0; q < 2 * TX; ++q) yn[q] = buf_load<T>(rs, voff, lane_u32(r, tile_rec_yhalo(NS, TY, q)) * cb);
    // lane q of moment vector v holds dword q % 8 of the moments of slot
    // 8v + q / 8 (slots the octant does not sweep carry cell 0: a harmless read)
#pragma unroll
    for (int v = 0; v < NM; ++v) {
      const int slot = 8 * v + (lane >> 3);
      const uint32_t cw = (uint32_t)__shfl((int)r, tile_rec_cell(NS, TX, TY, slot < NS ? slot : 0), kWave);
      mn[v] = momw[(size_t)(cw & 0x0fffffffu) * 8 + (lane & 7)];
    }
  };
  // records of the planes past the last one are read too (zeros, or the last
  // plane's record again): the loads they feed are harmless and keep the
  // loop's memory instructions unconditional
  const int last = a.tz;
  uint32_t rcur = rec[W + lane];
  issue(rcur);
  uint32_t rnext = rec[(size_t)(np > 1 ? 2 : 1) * W + lane];
  for (int c = 0; c < np; ++c) {
    T cc[NS], xh[2 * TY], yh[2 * TX];
    uint32_t mm[NM];
#pragma unroll
    for (int s = 0; s < NS; ++s) cc[s] = cn[s];
#pragma unroll
    for (int q = 0; q < 2 * TY; ++q) xh[q] = xn[q];
#pragma unroll
    for (int q = 0; q < 2 * TX; ++q) yh[q] = yn[q];
#pragma unroll
    for (int v = 0; v < NM; ++v) mm[v] = mn[v];
    const uint32_t r = rcur;
    issue(rnext);
    rcur = rnext;
    rnext = rec[(size_t)(c + 3 < last ? c + 3 : last) * W + lane];
    const uint32_t mask = lane_u32(r, 0);
#pragma unroll
    for (int b = 0; b < TY; ++b)
#pragma unroll
      for (int x = 0; x < TX; ++x) {
        const int s = b * TX + x;
        const bool sw = (mask >> s) & 1u;
        const uint32_t cw = lane_u32(r, tile_rec_cell(NS, TX, TY, s));
        const uint32_t cell = cw & 0x0fffffffu, flg = cw >> 28;
        double zz = 0.0;
        if (sw) {
          double nb[6];
          nb[0] = x >= 1 ? (double)cc[s - 1] : (double)xh[2 * b + 1];
          nb[1] = x >= 2 ? (double)cc[s - 2] : (double)xh[2 * b + x];
          nb[2] = b >= 1 ? (double)cc[s - TX] : (double)yh[2 * x + 1];
          nb[3] = b >= 2 ? (double)cc[s - 2 * TX] : (double)yh[2 * x + b];
          nb[4] = (double)p1[s];
          nb[5] = (double)p2[s];
          const double fp = (double)cc[s];
          const uint32_t mv = mm[s >> 3];
          const int ml = (s & 7) * 8;
          const double mr = lane_f64(mv, ml), mux = lane_f64(mv, ml + 2), muy = lane_f64(mv, ml + 4),
                       muz = lane_f64(mv, ml + 6);
          if (flg & 8u) {  // reads a wall slot the reference switched
            const ViewFix<T> vf = view_fix(a.view, cell, (int)o, ll, rhs);
#pragma unroll
            for (int q = 0; q < 4; ++q)
              if (vf.mask & (1u << q)) nb[2 + q] = (double)vf.v[q];
          }
          const bool fx = flg & 1u, fy = flg & 2u, fz = flg & 4u;
          const double mx = osx((int)o) > 0 ? mux : -mux;
          const double my = osy((int)o) > 0 ? muy : -muy;
          const double mz = osz((int)o) > 0 ? muz : -muz;
          double tt = cxl * mx;
          tt = tt + cyl * my;
          tt = tt + czl * mz;
          const double feq = wl * (mr + 1.0 * tt);
          double num = mu * (feq - 0.5 * fp);
          num = num + cxl * (fx ? a.d1 : a.c1) * nb[0];
          num = num + cxl * (fx ? 0.0 : a.c2) * nb[1];
          num = num + cyl * (fy ? a.d1 : a.c1) * nb[2];
          num = num + cyl * (fy ? 0.0 : a.c2) * nb[3];
          num = num + czl * (fz ? a.d1 : a.c1) * nb[4];
          num = num + czl * (fz ? 0.0 : a.c2) * nb[5];
          double den = den0 + cxl * (fx ? a.d1 : a.c0);
          den = den + cyl * (fy ? a.d1 : a.c0);
          den = den + czl * (fz ? a.d1 : a.c0);
          const T fnew = (T)(num / den);
          cc[s] = fnew;
          if (a.part) {
            const double fv = lact ? (double)fnew : 0.0;
            const double z = wave_sum4(fv, cxl * fv, cyl * fv, czl * fv);
            const uint32_t q = (uint32_t)lane >> 4;
            zz = ((kOctNeg >> (4 * o + q)) & 1u) ? -z : z;
          }
        }
        buf_store(cc[s], rs, voff_st, sw ? lane_u32(r, tile_rec_slot(s)) * cb : oob);
        buf_store(zz, rp, voff_p, sw ? ((cell - cell_lo) * 8u + o) * 32u : poob);
      }
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      p2[s] = p1[s];
      p1[s] = cc[s];
    }
  }
}

// ------------------------------------------------------------------ switch commit

// Final state of the switched wall slots after a sweep: the z switch if it
// fired, else the y switch (src/solver.F90:184-204 and analogues).
template <typename T>
__global__ void k_commit(T* f, const T* fw, const int32_t* walls, int nwalls,
                         const int32_t* wall_pos, const int32_t* wall_corner,
                         const uint64_t* views, Dims d) {
  f += blockIdx.y * d.rs_f;
  fw += blockIdx.y * d.rs_fw;
  const int lane = threadIdx.x & (kWave - 1);
  const int wpb = blockDim.x / kWave;
  // one wave per wall, striding over the list (the launch grid is capped)
  for (int idx = blockIdx.x * wpb + wave_in_block(); idx < nwalls; idx += gridDim.x * wpb) {
    const int w = walls[idx];
    const int32_t* wp = wall_pos + (size_t)w * Geometry::kWallPos;  // the wall cell per family
    const uint64_t vv = views[w];
    auto source = [&](int o, int src) -> const T* {
      return src <= 8 ? f + (size_t)(src - 1) * d.ostride + (size_t)wp[ofam(src - 1)] * d.Nc8
                      : fw + ((size_t)wall_corner[w] * 8 + o) * d.Nc8;
    };
    if (d.Nc8 <= kWave) {
      // all eight reads in flight before the first write: no switch source is
      // a switch target (checked statically in wallprog.h), so the reads see
      // the sweep's values whatever the order
      T v[8];
#pragma unroll
      for (int o = 0; o < 8; ++o) {
        const int src = (int)((vv >> (8 * o + 4)) & 15u);
        v[o] = (src != 0 && lane < d.Nc8) ? source(o, src)[lane] : T(0);
      }
#pragma unroll
      for (int o = 0; o < 8; ++o) {
        const int src = (int)((vv >> (8 * o + 4)) & 15u);
        if (src != 0 && lane < d.Nc8) f[(size_t)o * d.ostride + (size_t)wp[ofam(o)] * d.Nc8 + lane] = v[o];
      }
      continue;
    }
    for (int o = 0; o < 8; ++o) {
      const int src = (int)((vv >> (8 * o + 4)) & 15u);
      if (src == 0) continue;
      const T* s = source(o, src);
      T* dst = f + (size_t)o * d.ostride + (size_t)wp[ofam(o)] * d.Nc8;
      for (int l = lane; l < d.Nc8; l += kWave) dst[l] = s[l];
    }
  }
}

// ------------------------------------------------------------------ x exchange

// Ghost plane x = 0 receives plane x = Nx and ghost x = Nx + 1 receives plane
// x = 1, both as they were after the previous iteration's sweep (the
// reference's self send/receive posts last iteration's packed buffers,
// src/solver.F90:144-171, 535-560).  y/z ghost rows stay zero (MPI_PROC_NULL).
template <typename T>
__global__ void k_halo(T* f, T* buf_e, T* buf_w, Dims d) {
  const int64_t n = (int64_t)d.Ny * d.Nz * 8 * d.Nc8;
  f += blockIdx.y * d.rs_f;
  buf_e += blockIdx.y * d.rs_buf;
  buf_w += blockIdx.y * d.rs_buf;
  if (d.Nc8 <= kWave) {
    // one wave per (row, octant), one velocity per lane: the row's storage
    // positions are wave-uniform and all four reads are in flight together
    const int lane = threadIdx.x & (kWave - 1);
    const int wpb = blockDim.x / kWave;
    const int items = d.Ny * d.Nz * 8;
    for (int it = blockIdx.x * wpb + wave_in_block(); it < items; it += gridDim.x * wpb) {
      const int o = it & 7, r = it >> 3;
      const int j = 1 + r % d.Ny, k = 1 + r / d.Ny;
      const int64_t row = (int64_t)(j + 1) * d.Nxt + (int64_t)(k + 1) * d.Nxyt;
      if (lane >= d.Nc8) continue;
      const size_t t = (size_t)it * d.Nc8 + lane;
      const T ve = f[fat(d, o, row + d.Nx + 1) + lane], vw = f[fat(d, o, row + 2) + lane];
      const T be = buf_e[t], bw = buf_w[t];
      f[fat(d, o, row + 1) + lane] = be;
      f[fat(d, o, row + d.Nx + 2) + lane] = bw;
      buf_e[t] = ve;
      buf_w[t] = vw;
    }
    return;
  }
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int l = (int)(t % d.Nc8);
    int64_t r = t / d.Nc8;
    const int o = (int)(r % 8);
    r /= 8;
    const int j = 1 + (int)(r % d.Ny);
    const int k = 1 + (int)(r / d.Ny);
    const int64_t row = (int64_t)(j + 1) * d.Nxt + (int64_t)(k + 1) * d.Nxyt;
    const int64_t east = row + d.Nx + 1, west = row + 2;   // x = Nx, x = 1
    const int64_t g0 = row + 1, gN = row + d.Nx + 2;       // x = 0, x = Nx + 1
    const T ve = f[fat(d, o, east) + l], vw = f[fat(d, o, west) + l];
    f[fat(d, o, g0) + l] = buf_e[t];
    f[fat(d, o, gN) + l] = buf_w[t];
    buf_e[t] = ve;
    buf_w[t] = vw;
  }
}

// ------------------------------------------------------------------ wall BC

template <typename T>
struct WallArgs {
  T* f;
  T* fw;
  const int32_t* list;      // wall indices to process
  int n;
  const int32_t* wall_cell;
  const uint8_t* wall_label;
  const double* e1;
  const double* e2;
  const int32_t* wall_corner;
  const int32_t* wall_pos;  // [wall][Geometry::kWallPos] storage positions
  const double* cx;
  const double* cy;
  const double* cz;
  const double* w;
  double accom, inv_diff_flux, diff_flux;
  Dims d;
  int exact;                // sum the wall flux serially in the reference's order
  // coupled walls: per wall 6 snapshot indices (face, distance) or -1
  const int32_t* snap_idx;
  const T* snap;            // [snapshot][octant][velocity]
};

template <typename T, bool COUPLED>
__device__ __forceinline__ double load_nbr(const WallArgs<T>& a, int64_t cell, int o, int l,
                                           int slot) {
  if (COUPLED && slot >= 0) return (double)a.snap[((size_t)slot * 8 + o) * a.d.Nc8 + l];
  return (double)a.f[fat(a.d, o, cell) + l];
}

// src/solver.F90:772-1823: one wall, all velocities, on one wave.
template <int L, typename T, bool COUPLED>
__device__ void wall_one(const WallArgs<T>& a, int idx, int lane) {
  constexpr WallProg P = wall_prog(L);
  constexpr uint32_t written = wall_prog_written(P);
  const int wall = a.list[idx];
  const int64_t cell = a.wall_cell[wall];
  const int corner = a.wall_corner[wall];
  const double e1 = a.e1[wall], e2 = a.e2[wall];
  const int Nc8 = a.d.Nc8;
  const int64_t stride[3] = {1, a.d.Nxt, a.d.Nxyt};
  int sidx[6] = {-1, -1, -1, -1, -1, -1};
  if (COUPLED)
    for (int q = 0; q < 6; ++q) sidx[q] = a.snap_idx[idx * 6 + q];
  const double* cvel[3] = {a.cx, a.cy, a.cz};

  // wall flux per face: R_a = sum_l c_a(l) (v0 + v1 + v2 + v3) / DiffFlux
  double R[3] = {0.0, 0.0, 0.0};
  if (!a.exact && Nc8 <= kWave) {
    // one velocity per lane: every incoming value is loaded once, the fluxes
    // are summed from those values and the re-emission uses them; the same
    // operations in the same order as the two passes below
    const bool lact = lane < Nc8;
    const int l = lact ? lane : 0;
    // storage positions of the wall cell and of its neighbours (precomputed
    // per wall; coupled walls read some neighbours from the snapshot instead)
    const int32_t* wp = a.wall_pos + (size_t)wall * Geometry::kWallPos;
    const size_t os = (size_t)a.d.ostride;
    double v[24];
#pragma unroll
    for (int fa = 0; fa < P.nface; ++fa) {
      const int64_t st = P.face[fa].dir * stride[P.face[fa].axis];
      double s = 0.0;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int oo = P.face[fa].in[q];
        double n1, n2;
        if (COUPLED) {
          n1 = load_nbr<T, COUPLED>(a, cell + st, oo, l, sidx[fa * 2]);
          n2 = load_nbr<T, COUPLED>(a, cell + 2 * st, oo, l, sidx[fa * 2 + 1]);
        } else {
          n1 = (double)a.f[oo * os + (size_t)wp[4 + fa * 8 + q] * Nc8 + l];
          n2 = (double)a.f[oo * os + (size_t)wp[4 + fa * 8 + 4 + q] * Nc8 + l];
        }
        const double val = e1 * n1 + e2 * n2;
        v[P.face[fa].store + oo] = val;
        s = q == 0 ? val : s + val;
      }
      R[fa] = lact ? 0.0 + cvel[P.face[fa].axis][l] * s : 0.0;
    }
#pragma unroll
    for (int fa = 0; fa < P.nface; ++fa) R[fa] = wave_sum(R[fa]) / a.diff_flux;
    const double aw = a.accom * a.w[l], b = 1.0 - a.accom;
#pragma unroll
    for (int t = 0; t < P.nwr; ++t) v[P.wr[t].dst] = aw * R[P.wr[t].rho] + b * v[P.wr[t].src];
    if (!lact) return;
#pragma unroll
    for (int s = 0; s < 16; ++s) {
      if (!((written >> s) & 1u)) continue;
      if (s < 8)
        a.f[s * os + (size_t)wp[ofam(s)] * Nc8 + l] = (T)v[s];
      else
        a.fw[((size_t)corner * 8 + (s - 8)) * Nc8 + l] = (T)v[s];
    }
    return;
  }
  if (!a.exact) {
    for (int l = lane; l < Nc8; l += kWave) {
#pragma unroll
      for (int fa = 0; fa < P.nface; ++fa) {
        const int64_t st = P.face[fa].dir * stride[P.face[fa].axis];
        double s = 0.0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int oo = P.face[fa].in[q];
          const double val = e1 * load_nbr<T, COUPLED>(a, cell + st, oo, l, sidx[fa * 2]) +
                             e2 * load_nbr<T, COUPLED>(a, cell + 2 * st, oo, l, sidx[fa * 2 + 1]);
          s = q == 0 ? val : s + val;
        }
        R[fa] = R[fa] + cvel[P.face[fa].axis][l] * s;
      }
    }
#pragma unroll
    for (int fa = 0; fa < P.nface; ++fa) R[fa] = wave_sum(R[fa]) / a.diff_flux;
  } else {
    if (lane == 0) {
      for (int l = 0; l < Nc8; ++l) {
#pragma unroll
        for (int fa = 0; fa < P.nface; ++fa) {
          const int64_t st = P.face[fa].dir * stride[P.face[fa].axis];
          double s = 0.0;
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const int oo = P.face[fa].in[q];
            const double val = e1 * load_nbr<T, COUPLED>(a, cell + st, oo, l, sidx[fa * 2]) +
                               e2 * load_nbr<T, COUPLED>(a, cell + 2 * st, oo, l, sidx[fa * 2 + 1]);
            s = q == 0 ? val : s + val;
          }
          R[fa] = R[fa] + cvel[P.face[fa].axis][l] * s;
        }
      }
    }
#pragma unroll
    for (int fa = 0; fa < P.nface; ++fa) R[fa] = __shfl(R[fa], 0, kWave) / a.diff_flux;
  }

  // re-emission, in the reference's statement order
  for (int l = lane; l < Nc8; l += kWave) {
    double v[24];
#pragma unroll
    for (int fa = 0; fa < P.nface; ++fa) {
      const int64_t st = P.face[fa].dir * stride[P.face[fa].axis];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int oo = P.face[fa].in[q];
        v[P.face[fa].store + oo] = e1 * load_nbr<T, COUPLED>(a, cell + st, oo, l, sidx[fa * 2]) +
                                   e2 * load_nbr<T, COUPLED>(a, cell + 2 * st, oo, l, sidx[fa * 2 + 1]);
      }
    }
    const double aw = a.accom * a.w[l], b = 1.0 - a.accom;
#pragma unroll
    for (int t = 0; t < P.nwr; ++t) v[P.wr[t].dst] = aw * R[P.wr[t].rho] + b * v[P.wr[t].src];
#pragma unroll
    for (int s = 0; s < 16; ++s) {
      if (!((written >> s) & 1u)) continue;
      if (s < 8)
        a.f[fat(a.d, s, cell) + l] = (T)v[s];
      else
        a.fw[((size_t)corner * 8 + (s - 8)) * Nc8 + l] = (T)v[s];
    }
  }
}

// NF > 0: only the labels with NF faces are compiled in (a kernel for one
// face-count group of the uncoupled walls needs fewer registers)
template <typename T, bool COUPLED, int NF = 0>
__device__ void wall_dispatch(const WallArgs<T>& a, int idx, int lane) {
  switch (a.wall_label[a.list[idx]]) {
#define PK_CASE(L) \
  case L:          \
    if constexpr (NF == 0 || wall_prog(L).nface == NF) wall_one<L, T, COUPLED>(a, idx, lane); \
    break;
    PK_CASE(L_E) PK_CASE(L_W) PK_CASE(L_N) PK_CASE(L_S) PK_CASE(L_F) PK_CASE(L_B)
    PK_CASE(L_EN) PK_CASE(L_WN) PK_CASE(L_ES) PK_CASE(L_WS)
    PK_CASE(L_NF) PK_CASE(L_NB) PK_CASE(L_SF) PK_CASE(L_SB)
    PK_CASE(L_EF) PK_CASE(L_EB) PK_CASE(L_WF) PK_CASE(L_WB)
    PK_CASE(L_ENF) PK_CASE(L_WNF) PK_CASE(L_WSF) PK_CASE(L_ESF)
    PK_CASE(L_ENB) PK_CASE(L_WNB) PK_CASE(L_WSB) PK_CASE(L_ESB)
#undef PK_CASE
    default: break;
  }
}

// walls that read no other wall, all of one face count (NF): all at once,
// one wave each
template <typename T, int NF>
__global__ __launch_bounds__(256) void k_walls(WallArgs<T> a) {
  a.f += blockIdx.y * a.d.rs_f;
  a.fw += blockIdx.y * a.d.rs_fw;
  a.snap += blockIdx.y * a.d.rs_snap;
  const int wpb = blockDim.x / kWave;
  // one wave per wall, striding over the list (the launch grid is capped)
  for (int idx = blockIdx.x * wpb + wave_in_block(); idx < a.n; idx += gridDim.x * wpb)
    wall_dispatch<T, false, NF>(a, idx, threadIdx.x & (kWave - 1));
}

// walls that read another wall, one dependency level per launch: reads of
// earlier walls see their new values (lower levels ran before), reads of later
// walls come from the snapshot taken before the wall pass
template <typename T>
__global__ __launch_bounds__(256) void k_walls_coupled(WallArgs<T> a) {
  a.f += blockIdx.y * a.d.rs_f;
  a.fw += blockIdx.y * a.d.rs_fw;
  a.snap += blockIdx.y * a.d.rs_snap;
  const int wpb = blockDim.x / kWave;
  for (int idx = blockIdx.x * wpb + wave_in_block(); idx < a.n; idx += gridDim.x * wpb)
    wall_dispatch<T, true>(a, idx, threadIdx.x & (kWave - 1));
}

template <typename T>
__global__ void k_snapshot(const T* f, T* snap, const int32_t* cells, int n, Dims d) {
  const int64_t tot = (int64_t)n * 8 * d.Nc8;
  f += blockIdx.y * d.rs_f;
  snap += blockIdx.y * d.rs_snap;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < tot;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int l = (int)(t % d.Nc8);
    const int o = (int)((t / d.Nc8) % 8);
    const int s = (int)(t / (8 * d.Nc8));
    snap[t] = f[fat(d, o, cells[s]) + l];
  }
}

// ------------------------------------------------------------------ inlet/outlet, symmetry

// src/solver.F90:1832-1862: f(1) = f(0) + w dp for octants with cx > 0 and
// f(Nx) = f(Nx+1) - w dp for cx < 0, on every interior (j, k) row.
template <typename T>
__global__ void k_inout(T* f, const double* w, double pd, Dims d) {
  const int64_t n = (int64_t)d.Ny * d.Nz * 4 * d.Nc8;
  f += blockIdx.y * d.rs_f;
  if (d.Nc8 <= kWave) {
    // one wave per (row, octant pair), one velocity per lane; the two reads
    // (different octants, so different slots) are issued before the writes
    const int lane = threadIdx.x & (kWave - 1);
    const int wpb = blockDim.x / kWave;
    const int items = d.Ny * d.Nz * 4;
    for (int it = blockIdx.x * wpb + wave_in_block(); it < items; it += gridDim.x * wpb) {
      const int q = it & 3, r = it >> 2;
      const int j = 1 + r % d.Ny, k = 1 + r / d.Ny;
      const int64_t row = (int64_t)(j + 1) * d.Nxt + (int64_t)(k + 1) * d.Nxyt;
      const int oin = q == 0 ? 0 : q == 1 ? 3 : q == 2 ? 4 : 7;
      const int oout = q == 0 ? 1 : q == 1 ? 2 : q == 2 ? 5 : 6;
      if (lane >= d.Nc8) continue;
      const double wpd = w[lane] * pd;
      const double a = (double)f[fat(d, oin, row + 1) + lane];
      const double b = (double)f[fat(d, oout, row + d.Nx + 2) + lane];
      f[fat(d, oin, row + 2) + lane] = (T)(a + wpd);
      f[fat(d, oout, row + d.Nx + 1) + lane] = (T)(b - wpd);
    }
    return;
  }
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int l = (int)(t % d.Nc8);
    int64_t r = t / d.Nc8;
    const int q = (int)(r % 4);
    r /= 4;
    const int j = 1 + (int)(r % d.Ny);
    const int k = 1 + (int)(r / d.Ny);
    const int64_t row = (int64_t)(j + 1) * d.Nxt + (int64_t)(k + 1) * d.Nxyt;
    const int oin = q == 0 ? 0 : q == 1 ? 3 : q == 2 ? 4 : 7;
    const int oout = q == 0 ? 1 : q == 1 ? 2 : q == 2 ? 5 : 6;
    const double wpd = w[l] * pd;
    f[fat(d, oin, row + 2) + l] = (T)((double)f[fat(d, oin, row + 1) + l] + wpd);
    f[fat(d, oout, row + d.Nx + 1) + l] = (T)((double)f[fat(d, oout, row + d.Nx + 2) + l] - wpd);
  }
}

// src/solver.F90:1867-1892: y = 1 (south) then y = Ny (north) mirror copies.
template <typename T>
__device__ __forceinline__ void copy_slot(T* f, const Dims& d, int64_t cell, int l, int dst, int src) {
  f[fat(d, dst, cell) + l] = f[fat(d, src, cell) + l];
}
// four copies dst_k <- src_k in one cell, all reads before the writes (the
// four sources and the four targets are eight different octants)
template <typename T>
__device__ __forceinline__ void mirror4(T* f, const Dims& d, int64_t cell, int l, int d0, int s0, int d1,
                                        int s1, int d2, int s2, int d3, int s3) {
  const T v0 = f[fat(d, s0, cell) + l], v1 = f[fat(d, s1, cell) + l];
  const T v2 = f[fat(d, s2, cell) + l], v3 = f[fat(d, s3, cell) + l];
  f[fat(d, d0, cell) + l] = v0;
  f[fat(d, d1, cell) + l] = v1;
  f[fat(d, d2, cell) + l] = v2;
  f[fat(d, d3, cell) + l] = v3;
}

template <typename T>
__global__ void k_sym_y(T* f, Dims d) {
  const int64_t n = (int64_t)d.Nx * d.Nz * d.Nc8;
  f += blockIdx.y * d.rs_f;
  if (d.Nc8 <= kWave) {
    // one wave per (i, k), one velocity per lane: the four reads of a plane's
    // cell are issued before its four writes (sources and targets are
    // different octants); the south plane is done before the north one
    const int lane = threadIdx.x & (kWave - 1);
    const int wpb = blockDim.x / kWave;
    const int items = d.Nx * d.Nz;
    for (int it = blockIdx.x * wpb + wave_in_block(); it < items; it += gridDim.x * wpb) {
      const int i = 1 + it % d.Nx, k = 1 + it / d.Nx;
      if (lane >= d.Nc8) continue;
      const int64_t s = (int64_t)(i + 1) + 2 * (int64_t)d.Nxt + (int64_t)(k + 1) * d.Nxyt;
      mirror4<T>(f, d, s, lane, 1, 2, 0, 3, 5, 6, 4, 7);
      const int64_t q = (int64_t)(i + 1) + (int64_t)(d.Ny + 1) * d.Nxt + (int64_t)(k + 1) * d.Nxyt;
      mirror4<T>(f, d, q, lane, 2, 1, 3, 0, 6, 5, 7, 4);
    }
    return;
  }
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int l = (int)(t % d.Nc8);
    const int i = 1 + (int)((t / d.Nc8) % d.Nx);
    const int k = 1 + (int)(t / ((int64_t)d.Nc8 * d.Nx));
    const int64_t s = (int64_t)(i + 1) + 2 * (int64_t)d.Nxt + (int64_t)(k + 1) * d.Nxyt;
    copy_slot(f, d, s, l, 1, 2); copy_slot(f, d, s, l, 0, 3);
    copy_slot(f, d, s, l, 5, 6); copy_slot(f, d, s, l, 4, 7);
    const int64_t q = (int64_t)(i + 1) + (int64_t)(d.Ny + 1) * d.Nxt + (int64_t)(k + 1) * d.Nxyt;
    copy_slot(f, d, q, l, 2, 1); copy_slot(f, d, q, l, 3, 0);
    copy_slot(f, d, q, l, 6, 5); copy_slot(f, d, q, l, 7, 4);
  }
}

// src/solver.F90:1893-1918: z = 1 (back) then z = Nz (front).
template <typename T>
__global__ void k_sym_z(T* f, Dims d) {
  const int64_t n = (int64_t)d.Nx * d.Ny * d.Nc8;
  f += blockIdx.y * d.rs_f;
  if (d.Nc8 <= kWave) {
    // as k_sym_y: one wave per (i, j), back plane before front plane
    const int lane = threadIdx.x & (kWave - 1);
    const int wpb = blockDim.x / kWave;
    const int items = d.Nx * d.Ny;
    for (int it = blockIdx.x * wpb + wave_in_block(); it < items; it += gridDim.x * wpb) {
      const int i = 1 + it % d.Nx, j = 1 + it / d.Nx;
      if (lane >= d.Nc8) continue;
      const int64_t s = (int64_t)(i + 1) + (int64_t)(j + 1) * d.Nxt + 2 * (int64_t)d.Nxyt;
      mirror4<T>(f, d, s, lane, 0, 4, 1, 5, 2, 6, 3, 7);
      const int64_t q = (int64_t)(i + 1) + (int64_t)(j + 1) * d.Nxt + (int64_t)(d.Nz + 1) * d.Nxyt;
      mirror4<T>(f, d, q, lane, 4, 0, 5, 1, 6, 2, 7, 3);
    }
    return;
  }
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int l = (int)(t % d.Nc8);
    const int i = 1 + (int)((t / d.Nc8) % d.Nx);
    const int j = 1 + (int)(t / ((int64_t)d.Nc8 * d.Nx));
    const int64_t s = (int64_t)(i + 1) + (int64_t)(j + 1) * d.Nxt + 2 * (int64_t)d.Nxyt;
    copy_slot(f, d, s, l, 0, 4); copy_slot(f, d, s, l, 1, 5);
    copy_slot(f, d, s, l, 2, 6); copy_slot(f, d, s, l, 3, 7);
    const int64_t q = (int64_t)(i + 1) + (int64_t)(j + 1) * d.Nxt + (int64_t)(d.Nz + 1) * d.Nxyt;
    copy_slot(f, d, q, l, 4, 0); copy_slot(f, d, q, l, 5, 1);
    copy_slot(f, d, q, l, 6, 2); copy_slot(f, d, q, l, 7, 3);
  }
}

// The three plane kernels above in one launch: one wave per cell on the
// surface of the box [1, Nx] x [1, Ny] x [1, Nz] (the union of their ranges),
// one velocity per lane, applying in the reference's statement order the
// inlet or outlet (src/solver.F90:1832-1862), the south or north mirror
// (:1867-1892) and the back or front mirror (:1893-1918) that apply to the
// cell.  Each step reads and writes the cell itself (the inlet/outlet also
// reads the x ghost written by the exchange before), so the cells are
// independent and the result is that of the three kernels in sequence.
template <typename T>
__global__ __launch_bounds__(256) void k_planes(T* f, const double* w, double pd, const int32_t* cells, int64_t n,
                                                Dims d) {
  f += blockIdx.y * d.rs_f;
  const int lane = threadIdx.x & (kWave - 1);
  const int wpb = blockDim.x / kWave;
  for (int64_t it = (int64_t)blockIdx.x * wpb + wave_in_block(); it < n; it += (int64_t)gridDim.x * wpb) {
    const int64_t c = cells[it];
    const int k = (int)(c / d.Nxyt) - 1;
    const int j = (int)((c % d.Nxyt) / d.Nxt) - 1;
    const int i = (int)(c % d.Nxt) - 1;
    for (int l = lane; l < d.Nc8; l += kWave) {
      const double wpd = w[l] * pd;
      if (i == 1) {
        // inlet: octants with cx > 0 from the ghost x = 0
        const T a0 = f[fat(d, 0, c - 1) + l], a3 = f[fat(d, 3, c - 1) + l];
        const T a4 = f[fat(d, 4, c - 1) + l], a7 = f[fat(d, 7, c - 1) + l];
        f[fat(d, 0, c) + l] = (T)((double)a0 + wpd);
        f[fat(d, 3, c) + l] = (T)((double)a3 + wpd);
        f[fat(d, 4, c) + l] = (T)((double)a4 + wpd);
        f[fat(d, 7, c) + l] = (T)((double)a7 + wpd);
      }
      if (i == d.Nx) {
        // outlet: octants with cx < 0 from the ghost x = Nx + 1
        const T b1 = f[fat(d, 1, c + 1) + l], b2 = f[fat(d, 2, c + 1) + l];
        const T b5 = f[fat(d, 5, c + 1) + l], b6 = f[fat(d, 6, c + 1) + l];
        f[fat(d, 1, c) + l] = (T)((double)b1 - wpd);
        f[fat(d, 2, c) + l] = (T)((double)b2 - wpd);
        f[fat(d, 5, c) + l] = (T)((double)b5 - wpd);
        f[fat(d, 6, c) + l] = (T)((double)b6 - wpd);
      }
      if (j == 1) mirror4<T>(f, d, c, l, 1, 2, 0, 3, 5, 6, 4, 7);
      if (j == d.Ny) mirror4<T>(f, d, c, l, 2, 1, 3, 0, 6, 5, 7, 4);
      if (k == 1) mirror4<T>(f, d, c, l, 0, 4, 1, 5, 2, 6, 3, 7);
      if (k == d.Nz) mirror4<T>(f, d, c, l, 4, 0, 5, 1, 6, 2, 7, 3);
    }
  }
}

// ------------------------------------------------------------------ moments

// Interior fluid cells: their f is final after the sweep, so the moments are
// the sum of the eight octant partials.
__global__ void k_moments_partials(const double* part, const int32_t* cells, int64_t n, Mom* mom,
                                   int64_t ntotal) {
  part += blockIdx.y * ntotal * 32;
  mom += blockIdx.y * ntotal;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = cells[t];
    const double* p = part + (size_t)c * 32;
    double r = 0.0, x = 0.0, y = 0.0, z = 0.0;
    for (int q = 0; q < 8; ++q) {
      r += p[q * 4 + 0];
      x += p[q * 4 + 1];
      y += p[q * 4 + 2];
      z += p[q * 4 + 3];
    }
    mom[c] = Mom{r, x, y, z};
  }
}

// Moments straight from f (src/solver.F90:1925-1937), one wave per cell.
template <typename T>
__global__ __launch_bounds__(256) void k_moments_direct(const T* f, const int32_t* cells, int64_t n,
                                                        const double* cx, const double* cy,
                                                        const double* cz, Dims d, Mom* mom) {
  f += blockIdx.y * d.rs_f;
  mom += blockIdx.y * d.Ntotal;
  const int64_t idx = (int64_t)blockIdx.x * (blockDim.x / kWave) + wave_in_block();
  if (idx >= n) return;
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t c = cells[idx];
  size_t b8[8];
  for (int o = 0; o < 8; ++o) b8[o] = fat(d, o, c);
  double r = 0.0, x = 0.0, y = 0.0, z = 0.0;
  for (int l = lane; l < d.Nc8; l += kWave) {
    const double f1 = f[b8[0] + l], f2 = f[b8[1] + l], f3 = f[b8[2] + l], f4 = f[b8[3] + l];
    const double f5 = f[b8[4] + l], f6 = f[b8[5] + l], f7 = f[b8[6] + l], f8 = f[b8[7] + l];
    r += f1 + f2 + f3 + f4 + f5 + f6 + f7 + f8;
    x += cx[l] * (f1 - f2 - f3 + f4 + f5 - f6 - f7 + f8);
    y += cy[l] * (f1 + f2 - f3 - f4 + f5 + f6 - f7 - f8);
    z += cz[l] * (f1 + f2 + f3 + f4 - f5 - f6 - f7 - f8);
  }
  r = wave_sum(r); x = wave_sum(x); y = wave_sum(y); z = wave_sum(z);
  if (lane == 0) mom[c] = Mom{r, x, y, z};
}

// The same in the reference's summation order (one thread per cell), for
// bit-for-bit comparison with the oracle.
template <typename T>
__global__ void k_moments_serial(const T* f, const int32_t* cells, int64_t n, const double* cx,
                                 const double* cy, const double* cz, Dims d, Mom* mom) {
  f += blockIdx.y * d.rs_f;
  mom += blockIdx.y * d.Ntotal;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = cells[t];
    size_t b8[8];
    for (int o = 0; o < 8; ++o) b8[o] = fat(d, o, c);
    double r = 0.0, x = 0.0, y = 0.0, z = 0.0;
    for (int l = 0; l < d.Nc8; ++l) {
      const double f1 = f[b8[0] + l], f2 = f[b8[1] + l], f3 = f[b8[2] + l], f4 = f[b8[3] + l];
      const double f5 = f[b8[4] + l], f6 = f[b8[5] + l], f7 = f[b8[6] + l], f8 = f[b8[7] + l];
      r = r + f1 + f2 + f3 + f4 + f5 + f6 + f7 + f8;
      x = x + cx[l] * (f1 - f2 - f3 + f4 + f5 - f6 - f7 + f8);
      y = y + cy[l] * (f1 + f2 - f3 - f4 + f5 + f6 - f7 - f8);
      z = z + cz[l] * (f1 + f2 + f3 + f4 - f5 - f6 - f7 - f8);
    }
    mom[c] = Mom{r, x, y, z};
  }
}

// Ux on the plane x = column; with a pending defect correction (mom_e not
// null) the state's Ux is mom + mom_e.
__global__ void k_gather_plane(const Mom* mom, const Mom* mom_e, double* out, Dims d, int column) {
  const int64_t n = (int64_t)d.Ny * d.Nz;
  mom += blockIdx.y * d.Ntotal;
  if (mom_e) mom_e += blockIdx.y * d.Ntotal;
  out += blockIdx.y * n;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int j = 1 + (int)(t % d.Ny), k = 1 + (int)(t / d.Ny);
    const int64_t cl = (int64_t)(column + 1) + (int64_t)(j + 1) * d.Nxt + (int64_t)(k + 1) * d.Nxyt;
    out[t] = mom_e ? mom[cl].x + mom_e[cl].x : mom[cl].x;
  }
}

// ------------------------------------------------------------------ allocation-free mass flow

// Device scratch of piks_mass_flow_noalloc: the Ux plane x = 2 of right-hand
// side 0 (Ny * Nz doubles, at most kScratchDoubles).  One copy per device,
// held until the process ends; host code uses the address hipGetSymbolAddress
// returns for the current device, never this name.
constexpr size_t kScratchDoubles = (size_t)1 << 19;  // 4 MiB
static __device__ __align__(256) double scratch[kScratchDoubles];

// massLocal of src/output.F90:80-142 on one thread, in the reference's
// summation order (the same operations as ctx_mass on the host, so the same
// bits): u is the plane, j fastest; sq = sqrt(1/2), rl = 1 / Ref_L.
__global__ void k_mass_serial(const double* u, int Ny, int Nz, double ds, double sq, double pd, double rl,
                              double* out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  auto U = [&](int j, int k) { return u[(size_t)(j - 1) + (size_t)(k - 1) * Ny]; };
  double inner = 0.0, suth = 0.0, noth = 0.0, frnt = 0.0, back = 0.0;
  for (int k = 2; k <= Nz - 1; ++k)
    for (int j = 2; j <= Ny - 1; ++j) inner = inner + U(j, k) * ds * ds;
  for (int k = 2; k <= Nz - 1; ++k) suth = suth + U(1, k) * ds * ds * 0.5;
  for (int k = 2; k <= Nz - 1; ++k) noth = noth + U(Ny, k) * ds * ds * 0.5;
  for (int j = 2; j <= Ny - 1; ++j) back = back + U(j, 1) * ds * ds * 0.5;
  for (int j = 2; j <= Ny - 1; ++j) frnt = frnt + U(j, Nz) * ds * ds * 0.5;
  const double sb = U(1, 1) * ds * ds * 0.25, sf = U(1, Nz) * ds * ds * 0.25;
  const double nb = U(Ny, 1) * ds * ds * 0.25, nf = U(Ny, Nz) * ds * ds * 0.25;
  *out = (inner + suth + noth + frnt + back + sb + sf + nb + nf) * sq / pd / (rl * rl);
}

// ------------------------------------------------------------------ context

// ------------------------------------------------------------------ defect correction

// Element-wise steps of a defect-correction block (see Ctx) over arrays kept
// whole (fw, the exchange planes, the moments: every right-hand side at once).
// Before the block's FP64 iteration: s += e when a correction is pending,
// and old = s.
template <typename TS, typename TE>
__global__ void k_dc_begin_arr(TS* s, const TE* e, TS* old, size_t n, int fold) {
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
    const TS v = fold ? s[t] + (TS)e[t] : s[t];
    s[t] = v;
    old[t] = v;
  }
}
// After it: r = s - old.
template <typename TS, typename TR>
__global__ void k_dc_resid_arr(const TS* s, const TS* old, TR* r, size_t n) {
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x)
    r[t] = (TR)(s[t] - old[t]);
}
// Before each correction iteration: e = r (the block's first) or e + r.
template <typename TE, typename TR>
__global__ void k_dc_add_arr(TE* e, const TR* r, size_t n, int first) {
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x)
    e[t] = first ? (TE)r[t] : e[t] + (TE)r[t];
}
// Fold a pending correction into the state: s += e.
template <typename TS, typename TE>
__global__ void k_dc_flush_arr(TS* s, const TE* e, size_t n) {
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x)
    s[t] = s[t] + (TS)e[t];
}

// The same steps over the slots of f that no sweep updates (list entries
// storage index | octant << 28), one wave per slot and one velocity per lane;
// old holds [right-hand side][list entry][velocity].  The swept slots take
// these steps inside the sweep kernels (dc_previous, kDC).
// mode 0: begin (fold, copy), 1: residual, 2: add (first = e = r).
// With active lists (act_n not null; every mode launched with the same grid):
// mode 0 also clears the slot's correction, mode 1 records each slot whose
// residual has a nonzero bit in some velocity in the recording wave's own
// segment of act (positions wave + j * waves, j = 0, 1, ...; its count in
// act_n[wave], so no two waves share a counter), and mode 2 walks only those
// slots, each wave its own segment.  For a slot whose residual is zero
// throughout, e + r is e and the first correction's e = r is the cleared e,
// so skipping it changes no bit.
__global__ __launch_bounds__(256) void k_dc_slots(double* s, float* e, double* old, float* r,
                                                  const uint32_t* list, int64_t n, uint32_t* act,
                                                  uint32_t* act_n, Dims d, int mode, int flag) {
  const int rhs = blockIdx.y;
  s += rhs * d.rs_f;
  e += rhs * d.rs_f;
  r += rhs * d.rs_f;
  old += (size_t)rhs * n * d.Nc8;
  const int lane = threadIdx.x & (kWave - 1);
  const int wpb = blockDim.x / kWave;
  const int64_t gw = (int64_t)blockIdx.x * wpb + wave_in_block();
  const int64_t nw = (int64_t)gridDim.x * wpb;
  const bool use_act = act_n != nullptr;
  if (use_act) {
    act += (size_t)rhs * n;
    act_n += (size_t)rhs * nw;
  }
  if (use_act && mode == 2) {
    const int64_t cnt = (int64_t)act_n[gw];
    for (int64_t j = 0; j < cnt; ++j) {
      const uint32_t w = act[gw + j * nw];
      const size_t at = (size_t)(w >> 28) * d.ostride + (size_t)(w & 0x0fffffffu) * d.Nc8;
      for (int l = lane; l < d.Nc8; l += kWave) e[at + l] = flag ? r[at + l] : e[at + l] + r[at + l];
    }
    return;
  }
  uint32_t rec = 0;  // slots this wave recorded (mode 1)
  for (int64_t idx = gw; idx < n; idx += nw) {
    const uint32_t w = list[idx];
    const size_t at = (size_t)(w >> 28) * d.ostride + (size_t)(w & 0x0fffffffu) * d.Nc8;
    double* oo = old + (size_t)idx * d.Nc8;
    bool nz = false;
    for (int l = lane; l < d.Nc8; l += kWave) {
      if (mode == 0) {
        const double v = flag ? s[at + l] + (double)e[at + l] : s[at + l];
        s[at + l] = v;
        oo[l] = v;
        if (use_act) e[at + l] = 0.0f;
      } else if (mode == 1) {
        const float rv = (float)(s[at + l] - oo[l]);
        r[at + l] = rv;
        nz = nz || __builtin_bit_cast(uint32_t, rv) != 0u;
      } else {
        e[at + l] = flag ? r[at + l] : e[at + l] + r[at + l];
      }
    }
    if (use_act && mode == 1 && __builtin_amdgcn_ballot_w64(nz) != 0) {
      // rec counts this wave's earlier slots, so gw + rec * nw <= idx < n
      if (lane == 0) act[gw + (int64_t)rec * nw] = w;
      ++rec;
    }
  }
  if (use_act && mode == 1 && lane == 0) act_n[gw] = rec;
}

template <typename T>
int dalloc(Ctx& c, T** p, size_t n) {
  if (n == 0) n = 1;
  const size_t bytes = n * sizeof(T);
  hipError_t e = c.stream_ordered ? hipMallocAsync((void**)p, bytes, c.stream) : hipMalloc((void**)p, bytes);
  if (e != hipSuccess)
    return fail(PIKS_E_MEM, "device allocation of " + std::to_string(bytes) + " bytes: " + hipGetErrorString(e));
  c.allocs.push_back(*p);
  c.bytes += (int64_t)bytes;
  HIP_TRY(hipMemsetAsync(*p, 0, bytes, c.stream));
  return 0;
}

template <typename T>
int upload(Ctx& c, T** p, const std::vector<T>& v) {
  if (int rc = dalloc(c, p, v.size())) return rc;
  if (!v.empty()) HIP_TRY(hipMemcpyAsync(*p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, c.stream));
  return 0;
}

// Blocks for n work items of `block` per block, at most 65536: every kernel
// launched with it strides over its items.
int grid_for(int64_t n, int block) {
  int64_t g = (n + block - 1) / block;
  return (int)std::max<int64_t>(1, std::min<int64_t>(g, 65536));
}

int setup_velocity(Ctx& c) {
  const double *x0, *w0;
  const int n = c.prm.nc_fundamental;
  const bool exact = c.prm.exact_quadrature != 0;
  if (n < 1 || piks_gh_table(c.prm.half_range != 0, n, exact ? 1 : 0, &x0, &w0) != 0)
    return fail(PIKS_E_VGRID, "Nc_fundamental xi/wi for " + std::to_string(n) + " has not been provided");
  // src/velocityGrid.F90:166-190: nodes * sqrt(2), cx fastest, DiffFlux = 4 sum cz w
  std::vector<double> xi(n);
  for (int l = 0; l < n; ++l) xi[l] = x0[l] * std::sqrt(2.0);
  const int nc8 = n * n * n;
  c.cx.resize(nc8); c.cy.resize(nc8); c.cz.resize(nc8); c.w.resize(nc8);
  for (int l = 0; l < n; ++l)
    for (int m = 0; m < n; ++m)
      for (int q = 0; q < n; ++q) {
        const int k = q + m * n + l * n * n;
        c.cx[k] = xi[q]; c.cy[k] = xi[m]; c.cz[k] = xi[l];
        c.w[k] = w0[q] * w0[m] * w0[l];
      }
  double df = 0.0;
  for (int l = 0; l < nc8; ++l) df = df + c.cz[l] * c.w[l];
  c.diff_flux = df * 4;
  c.nc_fund = n;
  return 0;
}

template <typename T>
WallArgs<T> wall_args(Ctx& c, State<T>& st, const int32_t* list, int n) {
  WallArgs<T> a{};
  a.f = st.f; a.fw = st.fw; a.list = list; a.n = n;
  a.wall_cell = c.wall_cell; a.wall_label = c.wall_label; a.e1 = c.e1; a.e2 = c.e2;
  a.wall_corner = c.wall_corner;
  a.wall_pos = c.wall_pos;
  a.cx = c.dcx; a.cy = c.dcy; a.cz = c.dcz; a.w = c.dw;
  a.accom = c.prm.accom; a.diff_flux = c.diff_flux; a.inv_diff_flux = 1.0 / c.diff_flux;
  a.d = c.d; a.exact = c.exact;
  a.snap_idx = c.coupled_snap; a.snap = st.snap;
  return a;
}

// The sweep kernel for a workgroup size.
template <typename T>
using SweepKernel = void (*)(SweepArgs<T>);

template <typename T, int RB>
SweepKernel<T> sweep_kernel_rb(int block) {
  return block == 256 ? k_sweep<T, 256, false, RB> : k_sweep<T, 64, false, RB>;
}

// RB right-hand sides per wave (sharing the task record); with `recip` a
// single right-hand side multiplies by the tabulated 1 / denominator
template <typename T>
SweepKernel<T> sweep_kernel(int block, bool recip, int rb) {
  if (recip && rb == 1) return block == 256 ? k_sweep<T, 256, true, 1> : k_sweep<T, 64, true, 1>;
  switch (rb) {
    case 2: return sweep_kernel_rb<T, 2>(block);
    case 4: return sweep_kernel_rb<T, 4>(block);
    default: return sweep_kernel_rb<T, 1>(block);
  }
}

// The tiled sweep kernel for a tile cross-section (null: not compiled).
template <typename T>
void (*tile_kernel(int tx, int ty))(TileArgs<T>) {
  if (tx == 4 && ty == 4) return k_tile<T, 4, 4>;
  if (tx == 4 && ty == 2) return k_tile<T, 4, 2>;
  if (tx == 8 && ty == 2) return k_tile<T, 8, 2>;
  if (tx == 2 && ty == 2) return k_tile<T, 2, 2>;
  return nullptr;
}

// What one call of enqueue_t iterates: the moments it reads and writes, the
// inlet/outlet pressure drop, and the defect-correction role of its sweep
// (SweepArgs: dc 0 plain, 1 a block's FP64 iteration, 2 a correction
// iteration, with the fold / first flags).
struct IterMode {
  Mom* mom;
  double pd;
  int dc, fold, first;
};

// The single-solve sweep kernel: tasks per wave (1 with velocity passes).
template <typename T, bool kRecip, int kDC>
SweepKernel<T> sweep1_kernel(int nc8, int nt) {
  if (nc8 > kWave) return k_sweep1_passes<T, kRecip, kDC>;
  switch (nt) {
    case 2: return k_sweep1<T, kRecip, kDC, 2>;
    case 3: return k_sweep1<T, kRecip, kDC, 3>;
    case 4: return k_sweep1<T, kRecip, kDC, 4>;
    default: return k_sweep1<T, kRecip, kDC, 1>;
  }
}

// Tasks per wave of the single-solve sweep as launched.
inline int sweep1_tasks(const Ctx& c) { return c.d.Nc8 <= kWave ? c.sweep_nt : 1; }

// Several right-hand sides, one per wave: the single-solve task kernel with
// the right-hand side from blockIdx.y (grids of at most 64 velocities per
// octant, one-wave workgroups), unless the generic kernel is asked for.
inline bool multi_sweep1(const Ctx& c) {
  return c.nrhs > 1 && c.d.Nc8 <= kWave && c.sweep_block == 64 && !c.sweep_generic;
}

// Single-solve sweep kernels of a defect-correction role.
template <typename T, int kDC>
SweepKernel<T> sweep_kernel_dc(const Ctx& c, int rb) {
  if (c.nrhs == 1) {
    const bool recip = c.sweep_recip && !c.exact;
    return recip ? sweep1_kernel<T, true, kDC>(c.d.Nc8, c.sweep_nt) : sweep1_kernel<T, false, kDC>(c.d.Nc8, c.sweep_nt);
  }
  switch (rb) {
    case 2: return k_sweep<T, 64, false, 2, kDC>;
    case 4: return k_sweep<T, 64, false, 4, kDC>;
    default: return multi_sweep1(c) ? k_sweep1_rhs<T, kDC> : k_sweep<T, 64, false, 1, kDC>;
  }
}

// Enqueue nsteps iterations on the context's stream.  With ev0 >= 0, four
// events per step (steps ev0, ev0 + 1, ..., below kMaxTimedSteps) bracket the
// sweep, boundary and moment phases; the caller synchronises and reads them.
template <typename T>
int enqueue_t(Ctx& c, State<T>& st, int nsteps, int ev0, const IterMode& im) {
  const Geometry& g = c.g;
  const Dims d = c.d;
  SweepArgs<T> sa{};
  sa.f = st.f;
  sa.mom = im.mom;
  sa.dc_r = c.r32.f;
  sa.dc_e = c.s32.f;
  sa.dc_fold = im.fold;
  sa.dc_first = im.first;
  sa.part = c.exact ? nullptr : c.part;
  sa.cx = c.dcx; sa.cy = c.dcy; sa.cz = c.dcz; sa.w = c.dw;
  sa.mu = c.d_mu;
  sa.mu0 = c.mu_r[0];
  sa.inv_den = c.inv_den;
  sa.coef = c.coef;
  sa.c0 = 1.5 / g.ds; sa.c1 = 2.0 / g.ds; sa.c2 = -0.5 / g.ds; sa.d1 = 1.0 / g.ds;
  sa.view = static_cast<const ViewArgs<T>*>(c.view_args[sizeof(T) == sizeof(double) ? 0 : 1]);
  sa.Nc8 = d.Nc8;
  sa.nrhs = c.nrhs;
  sa.rs_f = d.rs_f;
  sa.rs_mom = d.Ntotal;
  // right-hand sides per wave of the several-Kn sweep (one record shared)
  const int rb_max = c.multi_rb > 0 ? c.multi_rb : 1;
  const int rb = c.nrhs >= 4 && rb_max >= 4 ? 4 : c.nrhs >= 2 && rb_max >= 2 ? 2 : 1;
  const unsigned ny_sweep = (unsigned)((c.nrhs + rb - 1) / rb), R = (unsigned)c.nrhs;
  const int wpb = 4;  // waves per block for one-wave-per-item kernels
  const int timed = ev0 >= 0 ? std::max(0, std::min(nsteps, kMaxTimedSteps - ev0)) : 0;
  while ((int)c.ev.size() < 4 * (ev0 + timed)) {
    hipEvent_t e;
    HIP_TRY(hipEventCreate(&e));
    c.ev.push_back(e);
  }
  SweepKernel<T> kdc = nullptr;
  if (im.dc != 0) {
    if constexpr (sizeof(T) == sizeof(double)) kdc = sweep_kernel_dc<T, 1>(c, rb);
    else kdc = sweep_kernel_dc<T, 2>(c, rb);
  }
  TileArgs<T> ta{};
  void (*ktile)(TileArgs<T>) = nullptr;
  if (c.sweep_tiles) {
    ta.f = st.f;
    ta.mom = im.mom;
    ta.part = c.exact ? nullptr : c.part;
    ta.cx = c.dcx; ta.cy = c.dcy; ta.cz = c.dcz; ta.w = c.dw;
    ta.mu = c.d_mu;
    ta.mu0 = c.mu_r[0];
    ta.view = sa.view;
    ta.c0 = sa.c0; ta.c1 = sa.c1; ta.c2 = sa.c2; ta.d1 = sa.d1;
    ta.rs_f = d.rs_f;
    ta.rs_mom = d.Ntotal;
    ta.ostride = d.ostride;
    ta.nrhs = c.nrhs;
    ta.Nc8 = d.Nc8;
    ta.tz = g.tl_tz;
    ktile = tile_kernel<T>(g.tl_tx, g.tl_ty);
    if (!ktile) return fail(PIKS_E_ARG, "no tiled sweep kernel for this tile shape");
  }
  for (int step = 0; step < nsteps; ++step) {
    const bool tm = step < timed;
    const int ev = 4 * (ev0 + step);
    if (tm) HIP_TRY(hipEventRecord(c.ev[ev + 0], c.stream));
    if (c.sweep_tiles) {
      const size_t per_tile = (size_t)(g.tl_tz + 1) * Geometry::kTileRecDw;
      for (int dg = 0; dg < g.tl_ndiag; ++dg) {
        const int64_t n = g.tl_off[(size_t)dg + 1] - g.tl_off[(size_t)dg];
        if (n == 0) continue;
        ta.ntile = (int32_t)n;
        ta.rec = c.tl_rec + (size_t)g.tl_off[(size_t)dg] * per_tile;
        const int64_t blocks = (n + 7) / 8 * 8;
        ta.xcd_per = c.xcd_remap ? (uint32_t)(blocks / 8) : 0u;
        hipLaunchKernelGGL(ktile, dim3((unsigned)blocks, R), dim3(64), 0, c.stream, ta);
      }
    }
    // Octant group gr sweeps octants [8 gr / ng, 8 (gr + 1) / ng) through
    // all hyperplanes (the octants' sweeps are independent).  With side
    // streams the groups' launch chains run concurrently, forked from and
    // joined back into the context's stream: one chain's launches fill the
    // GPU while another's drain and start.  Without them the groups run one
    // after the other.
    const bool flow = c.sweep_flow && sizeof(T) == sizeof(double) && im.dc == 0 && c.nrhs == 1;
    if (flow) {
      FlowArgs fa{};
      fa.f = reinterpret_cast<double*>(st.f);
      fa.mom = im.mom;
      fa.part = sa.part;
      fa.rec = c.flow_rec;
      fa.flags = c.flow_flags;
      fa.epoch = c.flow_epoch;
      fa.err = c.flow_err;
      fa.wave_ctr = c.flow_epoch + 8;
      fa.done = c.flow_epoch + 16;
      fa.mode = c.flow_epoch + 32;
      fa.next = c.flow_epoch + 40;
      fa.waves = (uint32_t)c.flow_waves;
      fa.barrier = (uint32_t)c.flow_barrier;
      fa.cx = c.dcx; fa.cy = c.dcy; fa.cz = c.dcz; fa.w = c.dw;
      fa.coef = c.coef;
      fa.fw = reinterpret_cast<const double*>(st.fw);
      fa.mu0 = c.mu_r[0];
      for (int q = 0; q < 9; ++q) fa.off[q] = c.flow_off[q];
      fa.ostride = d.ostride;
      fa.nslot = g.nstore + 1;
      fa.spin = 1u << 20;  // about 1-2 s of polls 16 x 64 cycles apart, far beyond any wait of a sweep
      fa.diag = c.flow_diag;
      FlowCounts fc{};
      for (int q = 0; q < 8; ++q) fc.n[q] = (uint32_t)(c.flow_off[q + 1] - c.flow_off[q]);
      hipLaunchKernelGGL(k_flow_epoch, dim3(1), dim3(64), 0, c.stream, c.flow_epoch);
      hipLaunchKernelGGL(c.flow_early ? k_sweep_flow<true> : k_sweep_flow<false>, dim3(8u * (unsigned)c.flow_waves),
                         dim3(64), 0, c.stream, fa, fa.rec,
                         fa.mom, fa.coef);
      hipLaunchKernelGGL(k_flow_done, dim3(1), dim3(64), 0, c.stream, c.flow_epoch, fc);
    }
    const int ng = c.sweep_tiles || flow ? 0 : std::max(1, c.oct_groups);
    const bool conc = ng > 1 && c.side[0] != nullptr;
    if (conc) {
      HIP_TRY(hipEventRecord(c.fork_ev, c.stream));
      for (int k = 1; k < ng; ++k) HIP_TRY(hipStreamWaitEvent(c.side[k - 1], c.fork_ev, 0));
    }
    for (int hg = 0; hg < ng * g.nhp; ++hg) {
      // concurrent groups are submitted hyperplane by hyperplane
      const int gr = conc ? hg % ng : hg / g.nhp, h = conc ? hg / ng : hg % g.nhp;
      const hipStream_t sst = conc && gr > 0 ? c.side[gr - 1] : c.stream;
      const int o0 = 8 * gr / ng, o1 = 8 * (gr + 1) / ng;
      const int64_t t0 = g.hp_oct[(size_t)h * 8 + o0], t1 = g.hp_oct[(size_t)h * 8 + o1];
      sa.ntask = (int32_t)(t1 - t0);
      if (sa.ntask == 0) continue;
      sa.tasks = reinterpret_cast<const TaskRec*>(c.hp_task) + t0;
      sa.win = c.hp_win + (size_t)h * 8;
      const int64_t wpg = c.sweep_block / kWave;
      const int64_t blocks = ((sa.ntask + wpg - 1) / wpg + 7) / 8 * 8;
      sa.xcd_per = c.xcd_remap ? (uint32_t)(blocks / 8) : 0u;
      // single solve: one-wave blocks of sweep1_tasks(c) tasks each
      const int64_t nt1 = sweep1_tasks(c);
      const int64_t blocks1 = ((sa.ntask + nt1 - 1) / nt1 + 7) / 8 * 8;
      if (kdc) {
        if (c.nrhs == 1) {
          sa.xcd_per = c.xcd_remap ? (uint32_t)(blocks1 / 8) : 0u;
          hipLaunchKernelGGL(kdc, dim3((unsigned)blocks1), dim3(64), 0, sst, sa);
        } else {
          // the defect-correction kernels of several right-hand sides have
          // one-wave workgroups whatever the sweep block size: one per task
          const int64_t blocks64 = ((int64_t)sa.ntask + 7) / 8 * 8;
          sa.xcd_per = c.xcd_remap ? (uint32_t)(blocks64 / 8) : 0u;
          hipLaunchKernelGGL(kdc, dim3((unsigned)blocks64, ny_sweep), dim3(64), 0, sst, sa);
        }
        continue;
      }
      if (c.nrhs == 1 && c.sweep_block == 64 && !c.sweep_generic) {
        const bool recip = c.sweep_recip && !c.exact;
        SweepKernel<T> k1 = recip ? sweep1_kernel<T, true, 0>(d.Nc8, c.sweep_nt)
                                  : sweep1_kernel<T, false, 0>(d.Nc8, c.sweep_nt);
        sa.xcd_per = c.xcd_remap ? (uint32_t)(blocks1 / 8) : 0u;
        hipLaunchKernelGGL(k1, dim3((unsigned)blocks1), dim3(64), 0, sst, sa);
        continue;
      }
      const SweepKernel<T> km = rb == 1 && multi_sweep1(c) ? k_sweep1_rhs<T, 0>
                                                           : sweep_kernel<T>(c.sweep_block, c.sweep_recip && !c.exact, rb);
      hipLaunchKernelGGL(km, dim3((unsigned)blocks, ny_sweep),
                         dim3(c.sweep_block), 0, sst, sa);
    }
    if (conc) {
      for (int k = 1; k < ng; ++k) {
        HIP_TRY(hipEventRecord(c.join_ev[k - 1], c.side[k - 1]));
        HIP_TRY(hipStreamWaitEvent(c.stream, c.join_ev[k - 1], 0));
      }
    }
    if (tm) HIP_TRY(hipEventRecord(c.ev[ev + 1], c.stream));
    if (!g.commit_walls.empty())
      hipLaunchKernelGGL(k_commit<T>, dim3(grid_for((int64_t)g.commit_walls.size(), wpb), R),
                         dim3(wpb * kWave), 0, c.stream, st.f, (const T*)st.fw, c.commit_walls,
                         (int)g.commit_walls.size(), c.wall_pos, c.wall_corner, c.views, d);
    hipLaunchKernelGGL(k_halo<T>,
                       dim3(d.Nc8 <= kWave ? grid_for((int64_t)g.Ny * g.Nz * 8, wpb)
                                           : grid_for((int64_t)g.Ny * g.Nz * 8 * d.Nc8, 256), R),
                       dim3(256), 0, c.stream, st.f, st.buf_e, st.buf_w, d);
    if (!g.snap_cell.empty())
      hipLaunchKernelGGL(k_snapshot<T>, dim3(grid_for((int64_t)g.snap_cell.size() * 8 * d.Nc8, 256), R),
                         dim3(256), 0, c.stream, (const T*)st.f, st.snap, c.snap_cell,
                         (int)g.snap_cell.size(), d);
    for (int nf = 1; nf <= 3; ++nf) {
      const int b = g.uncoupled_off[nf - 1], n = g.uncoupled_off[nf] - b;
      if (n == 0) continue;
      WallArgs<T> wa = wall_args<T>(c, st, c.uncoupled + b, n);
      void (*kw)(WallArgs<T>) = nf == 1 ? k_walls<T, 1> : nf == 2 ? k_walls<T, 2> : k_walls<T, 3>;
      hipLaunchKernelGGL(kw, dim3(grid_for(n, wpb), R), dim3(wpb * kWave), 0, c.stream, wa);
    }
    for (size_t lv = 0; lv + 1 < g.coupled_level_off.size(); ++lv) {
      const int b = g.coupled_level_off[lv], n = g.coupled_level_off[lv + 1] - b;
      WallArgs<T> wa = wall_args<T>(c, st, c.coupled + b, n);
      wa.snap_idx = c.coupled_snap + (size_t)b * 6;
      hipLaunchKernelGGL(k_walls_coupled<T>, dim3(grid_for(n, wpb), R), dim3(wpb * kWave), 0,
                         c.stream, wa);
    }
    if (c.planes_fused) {
      hipLaunchKernelGGL(k_planes<T>, dim3(grid_for((int64_t)g.surface.size(), wpb), R), dim3(wpb * kWave), 0,
                         c.stream, st.f, (const double*)c.dw, im.pd, (const int32_t*)c.surface,
                         (int64_t)g.surface.size(), d);
    } else {
      hipLaunchKernelGGL(k_inout<T>, dim3(d.Nc8 <= kWave ? grid_for((int64_t)g.Ny * g.Nz * 4, wpb) : grid_for((int64_t)g.Ny * g.Nz * 4 * d.Nc8, 256), R), dim3(256),
                         0, c.stream, st.f, (const double*)c.dw, im.pd, d);
      hipLaunchKernelGGL(k_sym_y<T>, dim3(d.Nc8 <= kWave ? grid_for((int64_t)g.Nx * g.Nz, wpb) : grid_for((int64_t)g.Nx * g.Nz * d.Nc8, 256), R), dim3(256), 0,
                         c.stream, st.f, d);
      hipLaunchKernelGGL(k_sym_z<T>, dim3(d.Nc8 <= kWave ? grid_for((int64_t)g.Nx * g.Ny, wpb) : grid_for((int64_t)g.Nx * g.Ny * d.Nc8, 256), R), dim3(256), 0,
                         c.stream, st.f, d);
    }
    if (tm) HIP_TRY(hipEventRecord(c.ev[ev + 2], c.stream));
    if (c.exact) {
      const int64_t n = (int64_t)g.fluid_interior.size() + (int64_t)g.fluid_boundary.size();
      hipLaunchKernelGGL(k_moments_serial<T>, dim3(grid_for(n, 64), R), dim3(64), 0, c.stream,
                         (const T*)st.f, c.fluid_all, n, c.dcx, c.dcy, c.dcz, d, im.mom);
    } else {
      const int64_t ni = (int64_t)g.fluid_interior.size(), nbd = (int64_t)g.fluid_boundary.size();
      if (ni > 0)
        hipLaunchKernelGGL(k_moments_partials, dim3(grid_for(ni, 256), R), dim3(256), 0, c.stream,
                           (const double*)c.part, c.fluid_int, ni, im.mom, g.Ntotal);
      if (nbd > 0)
        hipLaunchKernelGGL(k_moments_direct<T>, dim3((unsigned)((nbd + wpb - 1) / wpb), R),
                           dim3(wpb * kWave), 0, c.stream, (const T*)st.f, c.fluid_bnd, nbd,
                           c.dcx, c.dcy, c.dcz, d, im.mom);
    }
    if (tm) HIP_TRY(hipEventRecord(c.ev[ev + 3], c.stream));
  }
  HIP_TRY(hipGetLastError());
  return 0;
}

// Waves of every k_dc_slots launch of a context (four per block): the same
// in every mode, as its active lists require.
int64_t dc_slot_waves(const Ctx& c) { return 4 * (int64_t)grid_for(std::max<int64_t>(c.n_ns, 1), 4); }

// Launch shape of the element-wise defect-correction kernels.
dim3 dc_grid(size_t n) { return dim3((unsigned)std::max<size_t>(1, std::min<size_t>((n + 255) / 256, 16384))); }

// The element-wise part of a defect-correction step over the arrays kept
// whole and the non-swept slots of f (see k_dc_begin_arr ... k_dc_slots).
int dc_arrays(Ctx& c, int mode, int flag) {
  const size_t R = (size_t)c.nrhs;
  const size_t nfw = R * c.n_fw, nbuf = R * c.n_buf, nmom = R * (size_t)c.g.Ntotal * 4;
  double* mom = reinterpret_cast<double*>(c.mom);
  double* mom_e = reinterpret_cast<double*>(c.mom_e);
  double* mom_old = reinterpret_cast<double*>(c.mom_old);
  double* mom_r = reinterpret_cast<double*>(c.mom_r);
  hipStream_t s = c.stream;
  if (mode == 0) {
    hipLaunchKernelGGL((k_dc_begin_arr<double, float>), dc_grid(nfw), dim3(256), 0, s, c.s64.fw, (const float*)c.s32.fw, c.fw_old, nfw, flag);
    hipLaunchKernelGGL((k_dc_begin_arr<double, float>), dc_grid(nbuf), dim3(256), 0, s, c.s64.buf_e, (const float*)c.s32.buf_e, c.bufe_old, nbuf, flag);
    hipLaunchKernelGGL((k_dc_begin_arr<double, float>), dc_grid(nbuf), dim3(256), 0, s, c.s64.buf_w, (const float*)c.s32.buf_w, c.bufw_old, nbuf, flag);
    hipLaunchKernelGGL((k_dc_begin_arr<double, double>), dc_grid(nmom), dim3(256), 0, s, mom, (const double*)mom_e, mom_old, nmom, flag);
  } else if (mode == 1) {
    hipLaunchKernelGGL((k_dc_resid_arr<double, float>), dc_grid(nfw), dim3(256), 0, s, (const double*)c.s64.fw, (const double*)c.fw_old, c.r32.fw, nfw);
    hipLaunchKernelGGL((k_dc_resid_arr<double, float>), dc_grid(nbuf), dim3(256), 0, s, (const double*)c.s64.buf_e, (const double*)c.bufe_old, c.r32.buf_e, nbuf);
    hipLaunchKernelGGL((k_dc_resid_arr<double, float>), dc_grid(nbuf), dim3(256), 0, s, (const double*)c.s64.buf_w, (const double*)c.bufw_old, c.r32.buf_w, nbuf);
    hipLaunchKernelGGL((k_dc_resid_arr<double, double>), dc_grid(nmom), dim3(256), 0, s, (const double*)mom, (const double*)mom_old, mom_r, nmom);
  } else {
    hipLaunchKernelGGL((k_dc_add_arr<float, float>), dc_grid(nfw), dim3(256), 0, s, c.s32.fw, (const float*)c.r32.fw, nfw, flag);
    hipLaunchKernelGGL((k_dc_add_arr<float, float>), dc_grid(nbuf), dim3(256), 0, s, c.s32.buf_e, (const float*)c.r32.buf_e, nbuf, flag);
    hipLaunchKernelGGL((k_dc_add_arr<float, float>), dc_grid(nbuf), dim3(256), 0, s, c.s32.buf_w, (const float*)c.r32.buf_w, nbuf, flag);
    hipLaunchKernelGGL((k_dc_add_arr<double, double>), dc_grid(nmom), dim3(256), 0, s, mom_e, (const double*)mom_r, nmom, flag);
  }
  if (c.n_ns > 0)
    hipLaunchKernelGGL(k_dc_slots, dim3((unsigned)(dc_slot_waves(c) / 4), (unsigned)c.nrhs), dim3(256), 0, s, c.s64.f, c.s32.f,
                       c.ns_old, c.r32.f, (const uint32_t*)c.ns_list, c.n_ns, c.ns_act,
                       c.dc_active ? c.ns_act_n : nullptr, c.d, mode, flag);
  HIP_TRY(hipGetLastError());
  return 0;
}

// One iteration of a kind: kIterPlain, or, in a refined context's
// defect-correction block (see Ctx), the block's FP64 iteration (without or
// with folding the previous block's correction) or a correction iteration
// (the block's first or a later one), timed at event step ev (ev < 0:
// untimed).
enum { kIterPlain = 0, kIterBlock = 1, kIterBlockFold = 2, kIterCorrFirst = 3, kIterCorr = 4, kIterKinds = 5 };

int iteration_body(Ctx& c, int kind, int ev) {
  switch (kind) {
    case kIterPlain:
      return enqueue_t<double>(c, c.s64, 1, ev, IterMode{c.mom, c.prm.press_drop, 0, 0, 0});
    case kIterBlock:
    case kIterBlockFold: {
      // keep the state, iterate in FP64, take the residual
      const int fold = kind == kIterBlockFold ? 1 : 0;
      if (int rc = dc_arrays(c, 0, fold)) return rc;
      if (int rc = enqueue_t<double>(c, c.s64, 1, ev, IterMode{c.mom, c.prm.press_drop, 1, fold, 0})) return rc;
      return dc_arrays(c, 1, 0);
    }
    default: {
      // e <- A(e + r), pressure drop 0
      const int first = kind == kIterCorrFirst ? 1 : 0;
      if (int rc = dc_arrays(c, 2, first)) return rc;
      return enqueue_t<float>(c, c.s32, 1, ev, IterMode{c.mom_e, 0.0, 2, 0, first});
    }
  }
}

// Capture one iteration of a kind as a graph (kernel launches only: no
// memset or allocation node).
int capture_graph(Ctx& c, int kind) {
  if (c.graph[kind]) return 0;
  hipGraph_t gr = nullptr;
  HIP_TRY(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
  const int rc = iteration_body(c, kind, -1);
  const hipError_t e = hipStreamEndCapture(c.stream, &gr);
  if (rc) {
    if (gr) (void)hipGraphDestroy(gr);
    return rc;
  }
  if (e != hipSuccess) return fail(PIKS_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  const hipError_t ei = hipGraphInstantiate(&c.graph[kind], gr, nullptr, nullptr, 0);
  (void)hipGraphDestroy(gr);
  if (ei != hipSuccess) return fail(PIKS_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
  return 0;
}

// One iteration of a kind: replayed from its graph when the context uses
// graphs and the iteration is untimed, else launched.
int run_iteration(Ctx& c, int kind, int ev) {
  if (c.use_graph && ev < 0) {
    if (int rc = capture_graph(c, kind)) return rc;
    HIP_TRY(hipGraphLaunch(c.graph[kind], c.stream));
    return 0;
  }
  return iteration_body(c, kind, ev);
}

// nsteps iterations of a refined context (defect correction, see Ctx),
// timed from event step ev0 on (ev0 < 0: untimed).
int dc_iterate(Ctx& c, int nsteps, int ev0) {
  for (int step = 0; step < nsteps; ++step) {
    const int ev = ev0 >= 0 ? ev0 + step : -1;
    if (c.dc_len <= 1) {
      // plain FP64 iterations
      if (int rc = ctx_dc_flush(c)) return rc;
      if (int rc = run_iteration(c, kIterPlain, ev)) return rc;
      continue;
    }
    if (c.dc_pos == 0 || c.dc_pos >= c.dc_len) {
      if (int rc = run_iteration(c, c.dc_pending ? kIterBlockFold : kIterBlock, ev)) return rc;
      c.dc_pending = false;
      c.dc_pos = 1;
      continue;
    }
    if (int rc = run_iteration(c, c.dc_pos == 1 ? kIterCorrFirst : kIterCorr, ev)) return rc;
    c.dc_pending = true;
    ++c.dc_inner;
    if (++c.dc_pos >= c.dc_len) c.dc_pos = 0;
  }
  return 0;
}

// Timed synchronous iterations (piks_iterate).
int iterate(Ctx& c, int nsteps) {
  HIP_TRY(hipSetDevice(c.device));
  int rc = c.phase == PIKS_PREC_FP32_REFINED
               ? dc_iterate(c, nsteps, 0)
               : enqueue_t<double>(c, c.s64, nsteps, 0, IterMode{c.mom, c.prm.press_drop, 0, 0, 0});
  if (rc) return rc;
  HIP_TRY(hipStreamSynchronize(c.stream));
  if ((rc = flow_check(c))) return rc;
  const int timed = std::min(nsteps, kMaxTimedSteps);
  if (timed > 0) {
    double acc[3] = {0, 0, 0};
    for (int st = 0; st < timed; ++st)
      for (int q = 0; q < 3; ++q) {
        float t = 0;
        HIP_TRY(hipEventElapsedTime(&t, c.ev[4 * st + q], c.ev[4 * st + q + 1]));
        acc[q] += t;
      }
    float all = 0;
    HIP_TRY(hipEventElapsedTime(&all, c.ev[0], c.ev[4 * (timed - 1) + 3]));
    c.last_ms[0] = acc[0]; c.last_ms[1] = acc[1]; c.last_ms[2] = acc[2]; c.last_ms[3] = all;
  }
  return 0;
}

}  // namespace

// ------------------------------------------------------------------ context API

namespace piks {

int ctx_create_host(const piks_params* p, const uint8_t* raw, int64_t nbytes, Ctx** out) {
  if (!p || !raw || !out) return fail(PIKS_E_ARG, "null argument");
  *out = nullptr;
  if (p->precision != PIKS_PREC_FP64 && p->precision != PIKS_PREC_FP32_REFINED)
    return fail(PIKS_E_ARG, "unknown precision mode " + std::to_string(p->precision));
  if (!(p->kn > 0.0)) return fail(PIKS_E_ARG, "Kn must be positive");
  if (p->chk_converge_step < 1) return fail(PIKS_E_ARG, "chkConvergeStep must be >= 1");
  std::unique_ptr<Ctx> c(new Ctx());
  c->prm = *p;
  c->exact = p->exact_order != 0;
  c->device = p->device;
  if (int rc = setup_velocity(*c)) return rc;
  std::string err;
  GeomInput gi{p->nx, p->ny, p->nz, p->fluid_layer, p->wall_ext_order};
  if (p->wall_ext_order < 1 || p->wall_ext_order > 3)
    std::fprintf(stderr, " Error: wallExtOrder wroond shoud be [1|2|3]\n");  // src/physicalGrid.F90:947
  if (int rc = build_geometry(gi, raw, nbytes, c->g, err)) return fail(rc, err);
  const Geometry& g = c->g;
  c->d = Dims{g.Nx, g.Ny, g.Nz, g.Nxt, g.Nxyt, g.Ntotal, (int)c->cx.size(), {}, 0};
  c->d.ostride = (g.nstore + 1) * (int64_t)c->d.Nc8;   // + the zero slot
  c->n_f = (size_t)c->d.ostride * 8;
  c->n_fw = (size_t)std::max(g.ncorner, 1) * 8 * c->d.Nc8;
  c->n_buf = (size_t)g.Ny * g.Nz * 8 * c->d.Nc8;
  c->n_snap = std::max<size_t>(g.snap_cell.size(), 1) * 8 * c->d.Nc8;
  if (const char* b = std::getenv("PIKS_XCD_REMAP")) c->xcd_remap = std::atoi(b) != 0;
  if (const char* b = std::getenv("PIKS_SWEEP_BLOCK")) {
    const int v = std::atoi(b);
    if (v != 64 && v != 256) return fail(PIKS_E_ARG, "PIKS_SWEEP_BLOCK must be 64 or 256");
    c->sweep_block = v;
  }

  if (const char* b = std::getenv("PIKS_SWEEP_TILES")) c->sweep_tiles = std::atoi(b) != 0;
  if (const char* b = std::getenv("PIKS_TILE")) {
    int tx = 0, ty = 0, tz = 0;
    if (std::sscanf(b, "%dx%dx%d", &tx, &ty, &tz) != 3 || !tile_kernel<double>(tx, ty) || tz < 2 || tz > 256)
      return fail(PIKS_E_ARG, "PIKS_TILE must be TXxTYxTZ with TXxTY one of 4x4, 4x2, 8x2, 2x2 and TZ in 2..256");
    c->tile_tx = tx; c->tile_ty = ty; c->tile_tz = tz;
  }
  if (const char* b = std::getenv("PIKS_GRAPH")) c->use_graph = std::atoi(b) != 0;
  if (const char* b = std::getenv("PIKS_PLANES_FUSED")) c->planes_fused = std::atoi(b) != 0;
  // the sweep schedule of the parameters (a tuning table's choice) takes
  // precedence over the environment
  if (p->sweep == 1) {
    c->sweep_tiles = false;
  } else if (p->sweep == 2) {
    c->sweep_tiles = true;
    if (p->tile != 0) {
      const int tx = p->tile & 255, ty = (p->tile >> 8) & 255, tz = (p->tile >> 16) & 0xffff;
      if (!tile_kernel<double>(tx, ty) || tz < 2 || tz > 256)
        return fail(PIKS_E_ARG, "tile must be tx | ty << 8 | tz << 16 with tx x ty one of 4x4, 4x2, 8x2, 2x2");
      c->tile_tx = tx; c->tile_ty = ty; c->tile_tz = tz;
    }
  } else if (p->sweep != 0) {
    return fail(PIKS_E_ARG, "sweep must be 0, 1 or 2");
  }
  if (const char* b = std::getenv("PIKS_DC_BLOCK")) {
    const int v = std::atoi(b);
    if (v < 1 || v > 100000) return fail(PIKS_E_ARG, "PIKS_DC_BLOCK must be 1..100000");
    c->dc_fixed = v;
  }
  if (const char* b = std::getenv("PIKS_SWEEP_RECIP")) c->sweep_recip = std::atoi(b) != 0;
  if (const char* b = std::getenv("PIKS_SWEEP_GENERIC")) c->sweep_generic = std::atoi(b) != 0;
  if (const char* b = std::getenv("PIKS_OCT_GROUPS")) {
    const int v = std::atoi(b);
    if (v != 1 && v != 2 && v != 4 && v != 8) return fail(PIKS_E_ARG, "PIKS_OCT_GROUPS must be 1, 2, 4 or 8");
    c->oct_groups = v;
  }
  if (const char* b = std::getenv("PIKS_OCT_SERIAL")) c->oct_serial = std::atoi(b) != 0;
  if (const char* b = std::getenv("PIKS_DC_ACTIVE")) c->dc_active = std::atoi(b) != 0;
  // the single-launch sweep: on where it applies unless a setting that
  // shapes the hyperplane sweep asks for that path; PIKS_SWEEP_FLOW=0 turns
  // it off, =1 makes a context it cannot serve an error
  const char* const shaping[] = {"PIKS_XCD_REMAP", "PIKS_SWEEP_BLOCK", "PIKS_SWEEP_TILES", "PIKS_SWEEP_GENERIC",
                                 "PIKS_OCT_GROUPS", "PIKS_OCT_SERIAL", "PIKS_SWEEP_NT", "PIKS_SWEEP_RECIP"};
  c->flow_req = p->sweep != 0 ? 0 : -1;
  for (const char* v : shaping)
    if (std::getenv(v)) c->flow_req = 0;
  if (const char* b = std::getenv("PIKS_SWEEP_FLOW")) c->flow_req = std::atoi(b) != 0 ? 1 : 0;
  if (const char* b = std::getenv("PIKS_FLOW_WAVES")) {
    const int v = std::atoi(b);
    if (v < 8 || v > 4096) return fail(PIKS_E_ARG, "PIKS_FLOW_WAVES must be 8..4096");
    c->flow_waves = v;
  }
  if (const char* b = std::getenv("PIKS_FLOW_BARRIER")) c->flow_barrier = std::atoi(b);
  if (const char* b = std::getenv("PIKS_FLOW_EARLY")) c->flow_early_req = std::atoi(b) != 0 ? 1 : 0;
  if (const char* b = std::getenv("PIKS_FLOW_STRIPS")) {
    c->flow_strips = std::atoi(b);
    if (c->flow_strips < 1 || c->flow_strips > 64) return fail(PIKS_E_ARG, "PIKS_FLOW_STRIPS must be 1..64");
  }
  if (const char* b = std::getenv("PIKS_MULTI_RB")) {
    const int v = std::atoi(b);
    if (v != 1 && v != 2 && v != 4) return fail(PIKS_E_ARG, "PIKS_MULTI_RB must be 1, 2 or 4");
    c->multi_rb = v;
  }
  if (const char* b = std::getenv("PIKS_SWEEP_NT")) {
    const int v = std::atoi(b);
    if (v < 1 || v > 4) return fail(PIKS_E_ARG, "PIKS_SWEEP_NT must be 1..4");
    c->sweep_nt = v;
  }
  c->mu = std::sqrt((std::atan(1.0) * 4.0) / 2.0) / p->kn;   // src/flow.F90:90
  c->kn_r.assign(1, p->kn);
  c->mu_r.assign(1, c->mu);
  c->mass_r.assign(1, 1.0);
  c->d.rs_f = (int64_t)c->n_f;
  c->d.rs_fw = (int64_t)c->n_fw;
  c->d.rs_buf = (int64_t)c->n_buf;
  c->d.rs_snap = (int64_t)c->n_snap;
  // the tiled sweep serves grids of at most 64 velocities per octant and the
  // generic kernel is a forced alternative
  if (c->d.Nc8 > kWave || c->sweep_generic || p->precision == PIKS_PREC_FP32_REFINED) c->sweep_tiles = false;
  if (p->precision == PIKS_PREC_FP32_REFINED) {
    c->ns_host = build_nonswept(c->g, host_threads());
    c->n_ns = (int64_t)c->ns_host.size();
    c->dc_len = c->dc_fixed > 0 ? c->dc_fixed : 1;
  }
  if (c->sweep_tiles) {
    Geometry& gt = c->g;
    if (!build_tile_schedule(gt, c->tile_tx, c->tile_ty, c->tile_tz, host_threads()))
      return fail(PIKS_E_ARG, "tile shape does not fit the tile record");
    // every buffer offset of a tile (bytes, the zero slot included, and the
    // partial moments of its cells) is 32-bit
    if ((gt.tl_maxwin + 1) * c->d.Nc8 * (int64_t)sizeof(double) >= ((int64_t)1 << 31) ||
        (gt.tl_maxspan + 1) * 256 >= ((int64_t)1 << 31)) {
      c->sweep_tiles = false;
      gt.tl_rec.clear();
      gt.tl_rec.shrink_to_fit();
    }
  }
  // the sweep addresses one hyperplane's storage window with 32-bit offsets
  int64_t maxwin = 0;
  for (int32_t v : g.hp_len) maxwin = std::max<int64_t>(maxwin, v);
  if (maxwin * c->d.Nc8 * (int64_t)sizeof(double) >= ((int64_t)1 << 31))
    return fail(PIKS_E_ARG, "a hyperplane's storage window exceeds 2 GiB (image or velocity grid too large)");
  // task records as the sweep reads them: element offsets inside the
  // (hyperplane, octant) window, the zero slot as the window's end
  make_window_relative(c->g, c->d.Nc8, host_threads());
  *out = c.release();
  return 0;
}

size_t ctx_device_bytes(const Ctx& c) {
  const Geometry& g = c.g;
  const bool refined = c.prm.precision == PIKS_PREC_FP32_REFINED;
  const size_t R = (size_t)c.nrhs;
  const size_t state64 = R * (c.n_f + c.n_fw + 2 * c.n_buf + c.n_snap) * sizeof(double);
  // refined: the FP32 correction (with the snapshot) and residual, the FP64
  // copies of the non-swept slots and small arrays, three moment arrays
  const size_t state32 = refined ? R * (2 * (c.n_f + c.n_fw + 2 * c.n_buf) + c.n_snap) * sizeof(float) +
                                       R * ((size_t)c.n_ns * c.d.Nc8 + c.n_fw + 2 * c.n_buf) * sizeof(double) +
                                       R * (size_t)g.Ntotal * 3 * sizeof(Mom) + (size_t)c.n_ns * 4 +
                                       R * ((size_t)c.n_ns + 262144) * 4
                                 : 0;
  const size_t per_cell = (size_t)g.Ntotal * (R * (sizeof(Mom) + 32 * sizeof(double)) + 1 + 4 + 4 * 4);
  const size_t walls = g.wall_cell.size() * (8 + 4 + 4 + 1 + 8 + 8) + g.wall_pos.size() * 4 +
                       g.uncoupled.size() * 4 +
                       g.coupled.size() * 4 * 7 + g.snap_cell.size() * 4 + g.commit_walls.size() * 4;
  const size_t fluid = (size_t)(g.fluid_interior.size() + g.fluid_boundary.size()) * 8 + g.surface.size() * 4;
  const size_t sched = c.sweep_tiles ? g.tl_rec.size() * 4 : g.hp_task.size() * 4 + g.hp_lo.size() * 16;
  return state64 + state32 + per_cell + walls + fluid + sched +
         R * ((size_t)g.Ny * g.Nz * 8 + 8) + 4 * c.cx.size() * 8 + 64 * 256;
}

// The single-launch sweep's schedule and state, refused unless the context
// can use it: FP64 single solve, 64 velocities per octant, its own stream,
// slot indices below 2^28, W waves per octant resident at once (probe
// launches of the kernel itself) and W blocks on each of the eight XCDs (a
// probe launch of the same grid reads each block's XCD id).  Records:
// octant o's tasks in hyperplane order; words 1..6 the upwind sources
// (flow_source): the neighbours' slot indices, with bit 31 set on the
// nearest swept neighbour along each axis (the one the task waits for), and
// reads of wall slots the reference switched resolved to their source.
int64_t tasks_per_hyperplane(const Geometry& g);

int flow_setup(Ctx& c) {
  const Geometry& g = c.g;
  const int64_t nc8 = c.d.Nc8, ns = g.nstore;
  if (c.nrhs != 1 || nc8 != 64 || c.stream_ordered || c.prm.precision != PIKS_PREC_FP64 || ns + 1 >= (1 << 28))
    return fail(PIKS_E_ARG, "PIKS_SWEEP_FLOW needs a single FP64 solve with 64 velocities per octant on a "
                            "context with its own stream");
  // waves per octant: all resident at once (one-wave workgroups).  The
  // occupancy query is only an upper bound (it allowed 7 waves per SIMD where
  // 6 were resident), so a probe launch of the sweep kernel itself, with the
  // same registers, checks that every block of the grid is resident at once,
  // and fewer waves are tried if not.
  // Two loop structures (k_sweep_flow<kEarly>), each with its own registers:
  // the early upwind loads pay where an octant's hyperplanes average at least
  // two tasks per wave (measured: 40^3 and 64^3 packs faster without, 100^3
  // and 200^3 with), so the choice follows W, and the probe is repeated for
  // the kernel that will run.
  using FlowKernel = void (*)(FlowArgs, const TaskRec*, const Mom*, const SweepCoef*);
  int per_cu = 0, dev = 0;
  HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_sweep_flow<false>, 64, 0));
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDevice(&dev));
  HIP_TRY(hipGetDeviceProperties(&prop, dev));
  const int64_t resident = (int64_t)per_cu * prop.multiProcessorCount;
  const bool forced = c.flow_waves != 0;
  if (!forced) c.flow_waves = (int)std::min<int64_t>(resident / 8, 1024);
  if (c.flow_waves < 1 || 8 * (int64_t)c.flow_waves > resident)
    return fail(PIKS_E_ARG, "PIKS_SWEEP_FLOW: " + std::to_string(c.flow_waves) + " waves per octant do not fit the " +
                                std::to_string(resident) + " resident waves");
  int rc = 0;
  if ((rc = dalloc(c, &c.flow_epoch, 64)) || (rc = dalloc(c, &c.flow_diag, 2 * 8 * (size_t)c.flow_waves)))
    return rc;
  // flow_epoch: the launch's state (k_flow_epoch)
  c.flow_err = c.flow_epoch + 1;
  auto probe_waves = [&](FlowKernel kern) -> int {
    for (int attempt = 0;; ++attempt) {
      const unsigned nbp = 8u * (unsigned)c.flow_waves;
      FlowArgs pa{};
      pa.done = c.flow_epoch + 16;
      pa.diag = c.flow_diag;
      pa.probe = 1u << 12;  // about 1.4 ms; a launch places its blocks within tens of microseconds
      std::vector<uint32_t> seen(nbp, 0);
      HIP_TRY(hipMemsetAsync(c.flow_epoch, 0, 64 * sizeof(uint32_t), c.stream));
      hipLaunchKernelGGL(kern, dim3(nbp), dim3(64), 0, c.stream, pa, nullptr, nullptr, nullptr);
      HIP_TRY(hipMemcpyAsync(seen.data(), c.flow_diag, nbp * sizeof(uint32_t), hipMemcpyDeviceToHost, c.stream));
      HIP_TRY(hipStreamSynchronize(c.stream));
      // a block that gave up saw the arrivals up to then; the first to give
      // up saw only blocks that were resident together (none had left yet),
      // so the smallest count is what fits at once
      uint32_t fit = 0xffffffffu;
      for (uint32_t v : seen) fit = std::min(fit, v);
      if (fit == 0xffffffffu) return 0;
      const int w = (int)(fit / 8) - 8 * attempt;
      if (forced || attempt == 3 || w < 8)
        return fail(PIKS_E_HIP, "single-launch sweep: " + std::to_string(nbp) + " one-wave blocks are not "
                                "resident at once (" + std::to_string(fit) + " were)");
      c.flow_waves = std::min(w, c.flow_waves - 8);
    }
  };
  if ((rc = probe_waves(k_sweep_flow<false>))) return rc;
  if (c.flow_early_req >= 0) c.flow_early = c.flow_early_req != 0;
  else c.flow_early = tasks_per_hyperplane(g) / 8 >= 2 * (int64_t)c.flow_waves;
  if (c.flow_early && (rc = probe_waves(k_sweep_flow<true>))) return rc;
  const unsigned nb = 8u * (unsigned)c.flow_waves;
  {
    uint32_t* probe = nullptr;
    HIP_TRY(hipMalloc(&probe, nb * sizeof(uint32_t)));
    std::vector<uint32_t> xcc(nb, 99u);
    hipLaunchKernelGGL(k_flow_probe, dim3(nb), dim3(64), 0, c.stream, probe, nb);
    hipError_t e = hipMemcpyAsync(xcc.data(), probe, nb * sizeof(uint32_t), hipMemcpyDeviceToHost, c.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c.stream);
    (void)hipFree(probe);
    if (e != hipSuccess) return fail(PIKS_E_HIP, std::string("flow probe: ") + hipGetErrorString(e));
    // the sweep needs W waves on every XCD (blocks are dealt to the XCDs in
    // turn, from a start that varies between launches)
    std::vector<unsigned> per(16, 0);
    for (unsigned b = 0; b < nb; ++b) ++per[xcc[b] & 15u];
    for (unsigned x = 0; x < 8; ++x)
      if (per[x] != (unsigned)c.flow_waves)
        return fail(PIKS_E_HIP, "single-launch sweep: XCD " + std::to_string(x) + " ran " + std::to_string(per[x]) +
                                    " of the probe's " + std::to_string(nb) + " blocks, not " +
                                    std::to_string(c.flow_waves));
  }
  std::vector<int32_t> rec;
  std::string err;
  if ((rc = build_flow_records(g, (int)nc8, rec, c.flow_off, err, c.flow_strips))) return fail(rc, err);
  static_assert(sizeof(TaskRec) == Geometry::kTaskWords * sizeof(int32_t), "one record per task");
  if ((rc = dalloc(c, &c.flow_rec, rec.size() / Geometry::kTaskWords))) return rc;
  HIP_TRY(hipMemcpyAsync(c.flow_rec, rec.data(), rec.size() * sizeof(int32_t), hipMemcpyHostToDevice, c.stream));
  if ((rc = dalloc(c, &c.flow_flags, (size_t)8 * (ns + 1)))) return rc;
  HIP_TRY(hipMemsetAsync(c.flow_flags, 0, (size_t)8 * (ns + 1) * sizeof(uint32_t), c.stream));
  HIP_TRY(hipMemsetAsync(c.flow_epoch, 0, 64 * sizeof(uint32_t), c.stream));
  HIP_TRY(hipStreamSynchronize(c.stream));  // the host vectors above end here
  return 0;
}

// A failed wait of the single-launch sweep (its err flag), after the
// context's stream has been synchronised: the iteration's result is not
// valid; the context goes back to hyperplane launches.
int flow_check(Ctx& c) {
  if (!c.sweep_flow) return 0;
  uint32_t err[6] = {0, 0, 0, 0, 0, 0};
  HIP_TRY(hipMemcpy(err, c.flow_err, sizeof(err), hipMemcpyDeviceToHost));
  if (!err[0]) return 0;
  c.sweep_flow = false;
  const size_t nb = 8 * (size_t)c.flow_waves;
  std::vector<uint32_t> dg(2 * nb);
  HIP_TRY(hipMemcpy(dg.data(), c.flow_diag, dg.size() * 4, hipMemcpyDeviceToHost));
  // diag: block b swept octant dg[b] and did dg[nb + b] tasks
  std::vector<uint64_t> per(8, 0), done(8, 0);
  for (size_t b = 0; b < nb; ++b) {
    ++per[dg[b] & 7u];
    done[dg[b] & 7u] += dg[nb + b];
  }
  std::string blocks, tasks;
  for (int o = 0; o < 8; ++o) {
    blocks += " " + std::to_string(per[(size_t)o]);
    tasks += " " + std::to_string(done[(size_t)o]) + "/" + std::to_string(c.flow_off[o + 1] - c.flow_off[o]);
  }
  const std::string what =
      err[3] == 0xffffffffu   ? "octant " + std::to_string(err[1]) + " did " + std::to_string(err[2]) + " of its " +
                                  std::to_string(err[4]) + " tasks"
      : err[3] == 0xfffffffeu ? "XCD " + std::to_string(err[1]) + " started wave " + std::to_string(err[2]) +
                                    " of " + std::to_string(err[4])
                              : "a dependency wait gave up (octant " + std::to_string(err[1]) + ", task " + std::to_string(err[2]) +
                ", slot " + std::to_string(err[3]) + ", epoch " + std::to_string(err[4]) + ", flag seen " +
                std::to_string(err[5]) + ")";
  return fail(PIKS_E_HIP, "single-launch sweep: " + what + "; blocks per octant" + blocks + "; tasks done" + tasks +
                              "; the iteration's result is not valid and the context uses hyperplane launches "
                              "from now on");
}

// Average sweep tasks (swept cell, octant) per oriented hyperplane.
int64_t tasks_per_hyperplane(const Geometry& g) {
  return g.nhp > 0 && !g.hp_off.empty() ? g.hp_off[(size_t)g.nhp] / g.nhp : 0;
}

int ctx_create_device(Ctx& c, hipStream_t stream) {
  HIP_TRY(hipSetDevice(c.device));
  if (stream) {
    c.stream = stream;
    c.own_stream = false;
    c.stream_ordered = true;
  } else {
    HIP_TRY(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
    c.own_stream = true;
    size_t free_b = 0, total_b = 0;
    HIP_TRY(hipMemGetInfo(&free_b, &total_b));
    const size_t need = ctx_device_bytes(c);
    if (need > free_b)
      return fail(PIKS_E_MEM, "problem needs about " + std::to_string(need >> 20) + " MiB of device memory, " +
                                  std::to_string(free_b >> 20) + " MiB free");
  }
  // Concurrent octant groups measured fastest at 4 on the 200^3 sphere pack
  // (49 000 tasks per hyperplane on average: 19.1 ms per FP64 sweep, 19.5
  // with 2 groups, 20.8 with 1, 22.7 with 8, more streams than the process's
  // hardware queues).  On sphere packs of side 48-128 four chains against one
  // took 2.05 / 1.24 ms per iteration at 3 500 tasks per hyperplane, 2.64 /
  // 2.02 at 5 800, 4.25 / 4.37 at 12 200 and 7.66 / 8.14 at 20 900: below
  // ~10 000 a launch's quarter does not fill the chip, so the default takes
  // 4 from 10 000 tasks (times right-hand sides) per hyperplane on.  A
  // caller's stream (the server's slots) keeps one chain, its concurrency
  // coming from the slots.
  if (c.oct_groups == 0) {
    const int64_t per_hp = tasks_per_hyperplane(c.g) * c.nrhs;
    c.oct_groups = !stream && per_hp >= 10000 ? 4 : 1;
  }
  if (c.oct_groups > 1 && !c.oct_serial) {
    // default priority, like the context's stream
    HIP_TRY(hipEventCreateWithFlags(&c.fork_ev, hipEventDisableTiming));
    for (int k = 0; k + 1 < c.oct_groups; ++k) {
      HIP_TRY(hipStreamCreateWithFlags(&c.side[k], hipStreamNonBlocking));
      HIP_TRY(hipEventCreateWithFlags(&c.join_ev[k], hipEventDisableTiming));
    }
  }
  const Geometry& g = c.g;
  const bool refined = c.prm.precision == PIKS_PREC_FP32_REFINED;
  int rc = 0;
  const size_t R = (size_t)c.nrhs;
  if ((rc = dalloc(c, &c.s64.f, R * c.n_f)) || (rc = dalloc(c, &c.s64.fw, R * c.n_fw)) ||
      (rc = dalloc(c, &c.s64.buf_e, R * c.n_buf)) || (rc = dalloc(c, &c.s64.buf_w, R * c.n_buf)) ||
      (rc = dalloc(c, &c.s64.snap, R * c.n_snap)))
    return rc;
  if (refined) {
    if ((rc = dalloc(c, &c.s32.f, R * c.n_f)) || (rc = dalloc(c, &c.s32.fw, R * c.n_fw)) ||
        (rc = dalloc(c, &c.s32.buf_e, R * c.n_buf)) || (rc = dalloc(c, &c.s32.buf_w, R * c.n_buf)) ||
        (rc = dalloc(c, &c.s32.snap, R * c.n_snap)) || (rc = dalloc(c, &c.r32.f, R * c.n_f)) ||
        (rc = dalloc(c, &c.r32.fw, R * c.n_fw)) || (rc = dalloc(c, &c.r32.buf_e, R * c.n_buf)) ||
        (rc = dalloc(c, &c.r32.buf_w, R * c.n_buf)) ||
        (rc = dalloc(c, &c.ns_old, R * (size_t)std::max<int64_t>(c.n_ns, 1) * c.d.Nc8)) ||
        (rc = dalloc(c, &c.fw_old, R * c.n_fw)) || (rc = dalloc(c, &c.bufe_old, R * c.n_buf)) ||
        (rc = dalloc(c, &c.bufw_old, R * c.n_buf)) || (rc = dalloc(c, &c.mom_e, R * (size_t)g.Ntotal)) ||
        (rc = dalloc(c, &c.mom_old, R * (size_t)g.Ntotal)) || (rc = dalloc(c, &c.mom_r, R * (size_t)g.Ntotal)) ||
        (rc = upload(c, &c.ns_list, c.ns_host)) ||
        (rc = dalloc(c, &c.ns_act, R * (size_t)std::max<int64_t>(c.n_ns, 1))) ||
        (rc = dalloc(c, &c.ns_act_n, R * (size_t)dc_slot_waves(c))))
      return rc;
    c.phase = PIKS_PREC_FP32_REFINED;
  }
  if ((rc = dalloc(c, &c.mom, R * (size_t)g.Ntotal))) return rc;
  if ((rc = dalloc(c, &c.part, R * (size_t)g.Ntotal * 32))) return rc;
  if ((rc = dalloc(c, &c.plane, R * (size_t)g.Ny * g.Nz))) return rc;
  if ((rc = upload(c, &c.d_mu, c.mu_r))) return rc;
  {
    // 1 / denominator of the update (src/solver.F90:206-214) per first-order
    // flags and velocity, for right-hand side 0
    const double c0 = 1.5 / g.ds, d1 = 1.0 / g.ds;
    const int nc8 = c.d.Nc8;
    std::vector<double> inv((size_t)8 * nc8);
    for (int fl = 0; fl < 8; ++fl)
      for (int l = 0; l < nc8; ++l) {
        double den = 0.5 * c.mu_r[0] + c.cx[(size_t)l] * ((fl & 1) ? d1 : c0);
        den = den + c.cy[(size_t)l] * ((fl & 2) ? d1 : c0);
        den = den + c.cz[(size_t)l] * ((fl & 4) ? d1 : c0);
        inv[(size_t)fl * nc8 + l] = 1.0 / den;
      }
    if ((rc = upload(c, &c.inv_den, inv))) return rc;
    const double c1 = 2.0 / g.ds, c2 = -0.5 / g.ds;
    std::vector<SweepCoef> cf(8);
    for (int fl = 0; fl < 8; ++fl) {
      const bool fx = fl & 1, fy = fl & 2, fz = fl & 4;
      cf[(size_t)fl] = SweepCoef{fx ? d1 : c1, fx ? 0.0 : c2, fy ? d1 : c1, fy ? 0.0 : c2,
                                 fz ? d1 : c1, fz ? 0.0 : c2, fx ? d1 : c0, fy ? d1 : c0, fz ? d1 : c0, 0.0};
    }
    if ((rc = upload(c, &c.coef, cf))) return rc;
  }
  if ((rc = upload(c, &c.dcx, c.cx)) || (rc = upload(c, &c.dcy, c.cy)) ||
      (rc = upload(c, &c.dcz, c.cz)) || (rc = upload(c, &c.dw, c.w)))
    return rc;
  if ((rc = upload(c, &c.lab, g.lab)) ||
      (rc = upload(c, &c.wid, g.wid)) || (rc = upload(c, &c.views, g.wall_views)) ||
      (rc = upload(c, &c.wall_cell, g.wall_cell)) || (rc = upload(c, &c.wall_corner, g.wall_corner)) ||
      (rc = upload(c, &c.wall_label, g.wall_label)) || (rc = upload(c, &c.e1, g.wall_e1)) ||
      (rc = upload(c, &c.e2, g.wall_e2)) || (rc = upload(c, &c.uncoupled, g.uncoupled)) || (rc = upload(c, &c.wall_pos, g.wall_pos)) ||
      (rc = upload(c, &c.coupled, g.coupled)) || (rc = upload(c, &c.coupled_snap, g.coupled_snap)) ||
      (rc = upload(c, &c.snap_cell, g.snap_cell)) || (rc = upload(c, &c.commit_walls, g.commit_walls)) ||
      (rc = upload(c, &c.fluid_int, g.fluid_interior)) || (rc = upload(c, &c.fluid_bnd, g.fluid_boundary)) ||
      (rc = upload(c, &c.surface, g.surface)))
    return rc;
  std::vector<int32_t> all(g.fluid_interior);
  all.insert(all.end(), g.fluid_boundary.begin(), g.fluid_boundary.end());
  if ((rc = upload(c, &c.fluid_all, all))) return rc;
  if (c.sweep_tiles) {
    if ((rc = upload(c, &c.tl_rec, g.tl_rec))) return rc;
  } else {
    if ((rc = upload(c, &c.hp_task, g.hp_task))) return rc;
    std::vector<SweepWin> win(g.hp_lo.size());
    for (size_t q = 0; q < win.size(); ++q) {
      const int o = (int)(q % 8);
      win[q].off = (int64_t)o * c.d.ostride + (int64_t)g.hp_lo[q] * c.d.Nc8;
      win[q].len = (uint32_t)((int64_t)g.hp_len[q] * c.d.Nc8);
      win[q].pad = 0;
    }
    if ((rc = upload(c, &c.hp_win, win))) return rc;
    if (c.flow_req != 0) {
      const int r = flow_setup(c);
      if (r == 0) c.sweep_flow = true;
      else if (c.flow_req == 1) return r;
      else set_error(0, "");  // automatic choice: the hyperplane launches serve, and no error stays behind
    }
  }
  for (int F = 0; F < 4; ++F) {
    if ((rc = upload(c, &c.pos[F], g.pos[F]))) return rc;
    c.d.pos[F] = c.pos[F];
  }
  {
    std::vector<ViewArgs<double>> v64(1, ViewArgs<double>{c.s64.f, c.s64.fw, c.lab, c.wid, c.views, c.wall_corner, c.d});
    ViewArgs<double>* p64 = nullptr;
    if ((rc = upload(c, &p64, v64))) return rc;
    c.view_args[0] = p64;
    if (refined) {
      std::vector<ViewArgs<float>> v32(1, ViewArgs<float>{c.s32.f, c.s32.fw, c.lab, c.wid, c.views, c.wall_corner, c.d});
      ViewArgs<float>* p32 = nullptr;
      if ((rc = upload(c, &p32, v32))) return rc;
      c.view_args[1] = p32;
    }
  }
  // the uploads read host vectors, some local to this function: wait for
  // them here so that no copy outlives its source, whatever HIP does with
  // pageable memory (on a caller's stream this waits for that stream only)
  HIP_TRY(hipStreamSynchronize(c.stream));
  return 0;
}

int ctx_enqueue(Ctx& c, int nsteps) {
  HIP_TRY(hipSetDevice(c.device));
  if (c.phase == PIKS_PREC_FP32_REFINED) return dc_iterate(c, nsteps, -1);
  if (c.use_graph) {
    for (int s = 0; s < nsteps; ++s)
      if (int rc = run_iteration(c, kIterPlain, -1)) return rc;
    return 0;
  }
  return enqueue_t<double>(c, c.s64, nsteps, -1, IterMode{c.mom, c.prm.press_drop, 0, 0, 0});
}

int ctx_dc_flush(Ctx& c) {
  if (!c.dc_pending) {
    c.dc_pos = 0;
    return 0;
  }
  HIP_TRY(hipSetDevice(c.device));
  const size_t R = (size_t)c.nrhs;
  hipStream_t s = c.stream;
  const size_t nf = R * c.n_f, nfw = R * c.n_fw, nbuf = R * c.n_buf, nmom = R * (size_t)c.g.Ntotal * 4;
  hipLaunchKernelGGL((k_dc_flush_arr<double, float>), dc_grid(nf), dim3(256), 0, s, c.s64.f, (const float*)c.s32.f, nf);
  hipLaunchKernelGGL((k_dc_flush_arr<double, float>), dc_grid(nfw), dim3(256), 0, s, c.s64.fw, (const float*)c.s32.fw, nfw);
  hipLaunchKernelGGL((k_dc_flush_arr<double, float>), dc_grid(nbuf), dim3(256), 0, s, c.s64.buf_e, (const float*)c.s32.buf_e, nbuf);
  hipLaunchKernelGGL((k_dc_flush_arr<double, float>), dc_grid(nbuf), dim3(256), 0, s, c.s64.buf_w, (const float*)c.s32.buf_w, nbuf);
  hipLaunchKernelGGL((k_dc_flush_arr<double, double>), dc_grid(nmom), dim3(256), 0, s, reinterpret_cast<double*>(c.mom),
                     (const double*)reinterpret_cast<double*>(c.mom_e), nmom);
  HIP_TRY(hipGetLastError());
  c.dc_pending = false;
  c.dc_pos = 0;
  return 0;
}

void ctx_dc_policy(Ctx& c, double err) {
  if (c.phase != PIKS_PREC_FP32_REFINED) return;
  if (c.dc_fixed > 0) {
    c.dc_len = c.dc_fixed;
    return;
  }
  // FP32 rounds a block's correction e (about m times the per-iteration
  // change err, relative to the state) once per correction iteration, so a
  // block adds a relative error of at most about eps32 * m^2 * err; m is
  // sqrt(0.1 / err) (at most ~6e-9 per block; measured errors on the
  // benchmark images stay two orders of magnitude below this bound), in
  // whole check intervals, at most 200.  Above err = 1e-2 (the first checks
  // of a solve) every iteration is FP64.
  if (!(err <= 1e-2)) {
    c.dc_len = 1;
    return;
  }
  const int chk = std::max(1, c.prm.chk_converge_step);
  int m = (int)std::floor(std::sqrt(0.1 / std::max(err, 1e-300)));
  m = std::min(200, std::max(m, 2));
  if (m >= chk) m = m / chk * chk;
  c.dc_len = m;
}

int ctx_enqueue_plane(Ctx& c, double* host_plane) {
  const Geometry& g = c.g;
  HIP_TRY(hipSetDevice(c.device));
  hipLaunchKernelGGL(k_gather_plane, dim3(grid_for((int64_t)g.Ny * g.Nz, 256), (unsigned)c.nrhs), dim3(256), 0, c.stream,
                     (const Mom*)c.mom, c.dc_pending ? (const Mom*)c.mom_e : nullptr, c.plane, c.d, 2);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(host_plane, c.plane, (size_t)g.Ny * g.Nz * c.nrhs * sizeof(double),
                         hipMemcpyDeviceToHost, c.stream));
  return 0;
}

double ctx_mass(const Ctx& c, const double* u) {
  const Geometry& g = c.g;
  auto U = [&](int j, int k) { return u[(size_t)(j - 1) + (size_t)(k - 1) * g.Ny]; };
  const double ds = g.ds;
  const int Ny = g.Ny, Nz = g.Nz;
  double inner = 0.0, suth = 0.0, noth = 0.0, frnt = 0.0, back = 0.0;
  for (int k = 2; k <= Nz - 1; ++k)
    for (int j = 2; j <= Ny - 1; ++j) inner = inner + U(j, k) * ds * ds;
  for (int k = 2; k <= Nz - 1; ++k) suth = suth + U(1, k) * ds * ds * 0.5;
  for (int k = 2; k <= Nz - 1; ++k) noth = noth + U(Ny, k) * ds * ds * 0.5;
  for (int j = 2; j <= Ny - 1; ++j) back = back + U(j, 1) * ds * ds * 0.5;
  for (int j = 2; j <= Ny - 1; ++j) frnt = frnt + U(j, Nz) * ds * ds * 0.5;
  const double sb = U(1, 1) * ds * ds * 0.25, sf = U(1, Nz) * ds * ds * 0.25;
  const double nb = U(Ny, 1) * ds * ds * 0.25, nf = U(Ny, Nz) * ds * ds * 0.25;
  const double rl = 1.0 / c.prm.ref_l;
  return (inner + suth + noth + frnt + back + sb + sf + nb + nf) * std::sqrt(1.0 / 2.0) /
         c.prm.press_drop / (rl * rl);
}

void ctx_check(Ctx& c, int r, double m, double out[3]) {
  const double error = std::fabs(1.0 - m / c.mass_r[(size_t)r]) / (double)c.prm.chk_converge_step;
  c.mass_r[(size_t)r] = m;
  if (r == 0) c.mass = m;
  out[0] = m;
  out[1] = m * c.kn_r[(size_t)r] * std::sqrt(4.0 / (std::atan(1.0) * 4.0));
  out[2] = error;
}

int ctx_set_rhs(Ctx& c, int nrhs, const double* kn) {
  if (nrhs < 1 || nrhs > 1024) return fail(PIKS_E_ARG, "number of right-hand sides must be 1..1024");
  if (c.s64.f) return fail(PIKS_E_STATE, "right-hand sides are set before the device set-up");
  c.nrhs = nrhs;
  c.kn_r.assign(kn, kn + nrhs);
  c.mu_r.resize((size_t)nrhs);
  c.mass_r.assign((size_t)nrhs, 1.0);
  for (int r = 0; r < nrhs; ++r) {
    if (!(kn[r] > 0.0)) return fail(PIKS_E_ARG, "Kn must be positive");
    c.mu_r[(size_t)r] = std::sqrt((std::atan(1.0) * 4.0) / 2.0) / kn[r];   // src/flow.F90:90
  }
  c.mu = c.mu_r[0];
  return 0;
}

int ctx_set_phase(Ctx& c, int phase) {
  if (phase == c.phase) return 0;
  if (phase == PIKS_PREC_FP32_REFINED && !c.r32.f)
    return fail(PIKS_E_STATE, "context was created without FP32 correction storage");
  HIP_TRY(hipSetDevice(c.device));
  if (int rc = ctx_dc_flush(c)) return rc;
  if (c.own_stream) HIP_TRY(hipStreamSynchronize(c.stream));
  c.phase = phase;
  return 0;
}

void ctx_destroy(Ctx* c) {
  if (!c) return;
  // teardown is best effort: every release is attempted whatever the others
  // return
  (void)hipSetDevice(c->device);
  for (auto& g : c->graph)
    if (g) (void)hipGraphExecDestroy(g);
  if (c->stream_ordered) {
    for (void* p : c->allocs) (void)hipFreeAsync(p, c->stream);
  } else {
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (void* p : c->allocs) (void)hipFree(p);
  }
  for (auto& e : c->ev) (void)hipEventDestroy(e);
  for (int k = 0; k < 7; ++k) {
    if (c->side[k]) (void)hipStreamDestroy(c->side[k]);
    if (c->join_ev[k]) (void)hipEventDestroy(c->join_ev[k]);
  }
  if (c->fork_ev) (void)hipEventDestroy(c->fork_ev);
  if (c->stream && c->own_stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

}  // namespace piks

// ------------------------------------------------------------------ C ABI

extern "C" {

int piks_create(const piks_params* p, const uint8_t* raw_image, int64_t image_bytes, void** ctx_out) {
  g_err.clear();
  if (!ctx_out) return fail(PIKS_E_ARG, "null argument");
  *ctx_out = nullptr;
  Ctx* c = nullptr;
  if (int rc = ctx_create_host(p, raw_image, image_bytes, &c)) return rc;
  if (int rc = ctx_create_device(*c, nullptr)) {
    ctx_destroy(c);
    return rc;
  }
  *ctx_out = c;
  return 0;
}

int piks_iterate(void* ctx, int32_t nsteps) {
  if (!ctx || nsteps < 0) return fail(PIKS_E_ARG, "bad context or step count");
  Ctx& c = *static_cast<Ctx*>(ctx);
  if (c.use_graph) {
    if (int rc = ctx_enqueue(c, nsteps)) return rc;
    HIP_TRY(hipStreamSynchronize(c.stream));
    return flow_check(c);
  }
  return iterate(c, nsteps);
}

int piks_set_precision(void* ctx, int32_t precision) {
  if (!ctx) return fail(PIKS_E_ARG, "bad context");
  if (precision != PIKS_PREC_FP64 && precision != PIKS_PREC_FP32_REFINED)
    return fail(PIKS_E_ARG, "unknown precision mode " + std::to_string(precision));
  return ctx_set_phase(*static_cast<Ctx*>(ctx), precision);
}

int piks_set_correction_block(void* ctx, int32_t m) {
  if (!ctx) return fail(PIKS_E_ARG, "bad context");
  Ctx& c = *static_cast<Ctx*>(ctx);
  if (m < 0 || m > 100000) return fail(PIKS_E_ARG, "block length must be 0..100000");
  if (!c.r32.f) return fail(PIKS_E_STATE, "context was created without FP32 correction storage");
  c.dc_fixed = m;
  if (m > 0) c.dc_len = m;
  return 0;
}

static int mass_flow_sync(Ctx& c, double* m) {
  const size_t np = (size_t)c.g.Ny * c.g.Nz;
  std::vector<double> u(np * c.nrhs);
  if (int rc = ctx_enqueue_plane(c, u.data())) return rc;
  HIP_TRY(hipStreamSynchronize(c.stream));
  for (int r = 0; r < c.nrhs; ++r) m[r] = ctx_mass(c, u.data() + np * r);
  return 0;
}

// Per-device address of `scratch`, taken once per device with that device
// current (the array's host-side name is not a device address).
static double* scratch_on(int device) {
  static std::mutex mu;
  static double* ptr[256] = {};
  static bool done[256] = {};
  if (device < 0 || device >= 256) return nullptr;
  std::lock_guard<std::mutex> lk(mu);
  if (!done[device]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess) return nullptr;
    ptr[device] = static_cast<double*>(p);
    done[device] = true;
  }
  return ptr[device];
}

int piks_mass_flow_noalloc(void* ctx, double* mass_dev) {
  if (!ctx || !mass_dev) return fail(PIKS_E_ARG, "null argument");
  Ctx& c = *static_cast<Ctx*>(ctx);
  const Geometry& g = c.g;
  const size_t np = (size_t)g.Ny * g.Nz;
  if (np > kScratchDoubles)
    return fail(PIKS_E_ARG, "plane of " + std::to_string(np) + " cells exceeds the allocation-free limit of " +
                                std::to_string(kScratchDoubles) + " (use piks_mass_flow)");
  HIP_TRY(hipSetDevice(c.device));
  double* plane = scratch_on(c.device);
  if (!plane) return fail(PIKS_E_HIP, "hipGetSymbolAddress of the mass-flow scratch failed");
  hipLaunchKernelGGL(k_gather_plane, dim3(grid_for((int64_t)np, 256), 1u), dim3(256), 0, c.stream,
                     (const Mom*)c.mom, c.dc_pending ? (const Mom*)c.mom_e : nullptr, plane, c.d, 2);
  hipLaunchKernelGGL(k_mass_serial, dim3(1), dim3(64), 0, c.stream, (const double*)plane, g.Ny, g.Nz, g.ds,
                     std::sqrt(1.0 / 2.0), c.prm.press_drop, 1.0 / c.prm.ref_l, mass_dev);
  HIP_TRY(hipGetLastError());
  return 0;
}

int piks_mass_flow(void* ctx, double* m) {
  if (!ctx || !m) return fail(PIKS_E_ARG, "null argument");
  Ctx& c = *static_cast<Ctx*>(ctx);
  std::vector<double> all((size_t)c.nrhs);
  if (int rc = mass_flow_sync(c, all.data())) return rc;
  *m = all[0];
  return 0;
}

int piks_create_multi(const piks_params* p, const uint8_t* raw_image, int64_t image_bytes, int32_t nrhs,
                      const double* kn, void** ctx_out) {
  g_err.clear();
  if (!ctx_out || !kn) return fail(PIKS_E_ARG, "null argument");
  *ctx_out = nullptr;
  Ctx* c = nullptr;
  if (int rc = ctx_create_host(p, raw_image, image_bytes, &c)) return rc;
  int rc = ctx_set_rhs(*c, nrhs, kn);
  if (rc == 0) rc = ctx_create_device(*c, nullptr);
  if (rc) {
    ctx_destroy(c);
    return rc;
  }
  *ctx_out = c;
  return 0;
}

int piks_chk_converge_multi(void* ctx, double* out) {
  if (!ctx || !out) return fail(PIKS_E_ARG, "null argument");
  Ctx& c = *static_cast<Ctx*>(ctx);
  std::vector<double> m((size_t)c.nrhs);
  if (int rc = mass_flow_sync(c, m.data())) return rc;
  double worst = 0.0;
  for (int r = 0; r < c.nrhs; ++r) {
    ctx_check(c, r, m[(size_t)r], out + 3 * r);
    worst = std::isfinite(out[3 * r + 2]) ? std::max(worst, out[3 * r + 2]) : HUGE_VAL;
  }
  ctx_dc_policy(c, worst);
  return 0;
}

int piks_get_macro_rhs(void* ctx, int32_t rhs, double* rho, double* ux, double* uy, double* uz) {
  if (!ctx) return fail(PIKS_E_ARG, "null context");
  Ctx& c = *static_cast<Ctx*>(ctx);
  if (rhs < 0 || rhs >= c.nrhs) return fail(PIKS_E_ARG, "right-hand side out of range");
  HIP_TRY(hipSetDevice(c.device));
  if (int rc = ctx_dc_flush(c)) return rc;
  std::vector<Mom> h((size_t)c.g.Ntotal);
  HIP_TRY(hipMemcpyAsync(h.data(), c.mom + (size_t)rhs * c.g.Ntotal, h.size() * sizeof(Mom),
                         hipMemcpyDeviceToHost, c.stream));
  HIP_TRY(hipStreamSynchronize(c.stream));
  for (size_t q = 0; q < h.size(); ++q) {
    if (rho) rho[q] = h[q].r;
    if (ux) ux[q] = h[q].x;
    if (uy) uy[q] = h[q].y;
    if (uz) uz[q] = h[q].z;
  }
  return 0;
}

int piks_chk_converge(void* ctx, double out[3]) {
  if (!ctx || !out) return fail(PIKS_E_ARG, "null argument");
  Ctx& c = *static_cast<Ctx*>(ctx);
  // one mass flow per right-hand side (a multi context reports the first)
  std::vector<double> m((size_t)c.nrhs);
  if (int rc = mass_flow_sync(c, m.data())) return rc;
  ctx_check(c, 0, m[0], out);  // src/output.F90:151-155
  ctx_dc_policy(c, out[2]);
  return 0;
}

int piks_get_macro(void* ctx, double* rho, double* ux, double* uy, double* uz) {
  if (!ctx) return fail(PIKS_E_ARG, "null context");
  Ctx& c = *static_cast<Ctx*>(ctx);
  HIP_TRY(hipSetDevice(c.device));
  if (int rc = ctx_dc_flush(c)) return rc;
  std::vector<Mom> h((size_t)c.g.Ntotal);
  HIP_TRY(hipMemcpyAsync(h.data(), c.mom, h.size() * sizeof(Mom), hipMemcpyDeviceToHost, c.stream));
  HIP_TRY(hipStreamSynchronize(c.stream));
  for (size_t q = 0; q < h.size(); ++q) {
    if (rho) rho[q] = h[q].r;
    if (ux) ux[q] = h[q].x;
    if (uy) uy[q] = h[q].y;
    if (uz) uz[q] = h[q].z;
  }
  return 0;
}

int piks_get_f(void* ctx, double* f) {
  if (!ctx || !f) return fail(PIKS_E_ARG, "null argument");
  Ctx& c = *static_cast<Ctx*>(ctx);
  HIP_TRY(hipSetDevice(c.device));
  const size_t nc8 = c.d.Nc8, nt = (size_t)c.g.Ntotal;
  std::vector<double> h(c.n_f);
  if (int rc = ctx_dc_flush(c)) return rc;
  HIP_TRY(hipMemcpyAsync(h.data(), c.s64.f, h.size() * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  HIP_TRY(hipStreamSynchronize(c.stream));
  // device [o][storage index][l] -> Fortran f(cell, l, o)
  for (size_t o = 0; o < 8; ++o) {
    const std::vector<int32_t>& pos = c.g.pos[ofam((int)o)];
    for (size_t cell = 0; cell < nt; ++cell) {
      const double* src = &h[o * (size_t)c.d.ostride + (size_t)pos[cell] * nc8];
      for (size_t l = 0; l < nc8; ++l) f[(o * nc8 + l) * nt + cell] = src[l];
    }
  }
  return 0;
}

int piks_get_fw(void* ctx, double* fw) {
  if (!ctx || !fw) return fail(PIKS_E_ARG, "null argument");
  Ctx& c = *static_cast<Ctx*>(ctx);
  HIP_TRY(hipSetDevice(c.device));
  const size_t nc8 = c.d.Nc8, ncor = (size_t)c.g.ncorner;
  if (ncor == 0) return 0;
  std::vector<double> h(ncor * 8 * nc8);
  if (int rc = ctx_dc_flush(c)) return rc;
  HIP_TRY(hipMemcpyAsync(h.data(), c.s64.fw, h.size() * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  HIP_TRY(hipStreamSynchronize(c.stream));
  for (size_t o = 0; o < 8; ++o)
    for (size_t q = 0; q < ncor; ++q)
      for (size_t l = 0; l < nc8; ++l) fw[(o * nc8 + l) * ncor + q] = h[(q * 8 + o) * nc8 + l];
  return 0;
}

int piks_get_labels(void* ctx, int32_t* labels) {
  if (!ctx || !labels) return fail(PIKS_E_ARG, "null argument");
  Ctx& c = *static_cast<Ctx*>(ctx);
  for (size_t q = 0; q < c.g.lab.size(); ++q) labels[q] = c.g.lab[q];
  return 0;
}

int piks_get_info(void* ctx, piks_info* info) {
  if (!ctx || !info) return fail(PIKS_E_ARG, "null argument");
  Ctx& c = *static_cast<Ctx*>(ctx);
  const Geometry& g = c.g;
  std::memset(info, 0, sizeof(*info));
  info->nx = g.Nx; info->ny = g.Ny; info->nz = g.Nz;
  info->nxt = g.Nxt; info->nyt = g.Nyt; info->nzt = g.Nzt;
  info->ntotal = g.Ntotal;
  info->nc8 = c.d.Nc8;
  info->nwall = (int32_t)g.wall_cell.size();
  info->ncorner = g.ncorner;
  info->nfluid = g.nfluid;
  info->updates_per_iter = g.nfluid * 8 * (int64_t)c.d.Nc8;
  info->ds = g.ds; info->mu = c.mu; info->diff_flux = c.diff_flux;
  info->mass = c.mass;
  info->device_bytes = c.bytes;
  info->coupled_walls = (int32_t)g.coupled.size();
  info->precision = c.phase;
  info->steps_fp32 = c.dc_inner;
  info->dc_block = c.phase == PIKS_PREC_FP32_REFINED ? c.dc_len : 0;
  info->sweep_path = c.sweep_flow ? 2 : c.sweep_tiles ? 1 : 0;
  return 0;
}

int piks_preprocess(const piks_params* p, const uint8_t* raw_image, int64_t image_bytes,
                    int32_t* labels, piks_info* info) {
  g_err.clear();
  if (!p || !raw_image) return fail(PIKS_E_ARG, "null argument");
  Geometry g;
  std::string err;
  GeomInput gi{p->nx, p->ny, p->nz, p->fluid_layer, p->wall_ext_order};
  if (int rc = build_geometry(gi, raw_image, image_bytes, g, err)) return fail(rc, err);
  if (labels)
    for (size_t q = 0; q < g.lab.size(); ++q) labels[q] = g.lab[q];
  if (info) {
    std::memset(info, 0, sizeof(*info));
    info->nx = g.Nx; info->ny = g.Ny; info->nz = g.Nz;
    info->nxt = g.Nxt; info->nyt = g.Nyt; info->nzt = g.Nzt;
    info->ntotal = g.Ntotal;
    info->nwall = (int32_t)g.wall_cell.size();
    info->ncorner = g.ncorner;
    info->nfluid = g.nfluid;
    info->ds = g.ds;
    info->coupled_walls = (int32_t)g.coupled.size();
  }
  return 0;
}

int piks_last_timing(void* ctx, double ms[4]) {
  if (!ctx || !ms) return fail(PIKS_E_ARG, "null argument");
  Ctx& c = *static_cast<Ctx*>(ctx);
  for (int q = 0; q < 4; ++q) ms[q] = c.last_ms[q];
  return 0;
}

int piks_destroy(void* ctx) {
  ctx_destroy(static_cast<Ctx*>(ctx));
  return 0;
}

const char* piks_last_error(void) { return g_err.c_str(); }

}  // extern "C"
