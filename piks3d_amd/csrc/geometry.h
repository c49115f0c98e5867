// geometry.h -- host preprocessing of the voxel image into the device layouts.
//
// Follows src/physicalGrid.F90:217-1625 (single rank): read the 0/1 image into
// a grid with two ghost layers, drill one-voxel pores and remove one-voxel
// walls twice, label wall cells by their fluid sides, list walls in raster
// order with their extrapolation coefficients, and derive per-cell
// first-order flags of the upwind scheme.  Instead of the reference's eight
// per-octant raster lists and 9-coefficient arrays it produces what the GPU
// sweep needs: one byte label per cell, a 32-bit code per fluid cell (three
// first-order flags per octant plus a "reads a switched wall" flag per
// octant), and per wall the sources of the reference's in-sweep switches.
#pragma once
#include <stdint.h>

#include <algorithm>
#include <atomic>
#include <string>
#include <thread>
#include <vector>

namespace piks {

struct GeomInput {
  int nx, ny, nz;           // image size as in para.in (nx before fluid layers)
  int fluid_layer;
  int wall_ext_order;
  int threads = 0;          // host threads for the preprocessing (0: host_threads())
};

// Host threads the preprocessing of one image uses: the calling thread's
// setting if one was made (set_host_threads), else PIKS_HOST_THREADS from the
// environment, else the machine's hardware threads, at most 16.  Every
// parallel loop writes disjoint data, so the result does not depend on the
// count.
int host_threads();
void set_host_threads(int n);  // this thread only; 0 restores the default

// Run f(0..n-1) on up to `threads` threads (the calling thread included).
template <typename F>
void parallel_for(int n, int threads, F&& f);

struct Geometry {
  int Nx, Ny, Nz;           // after fluid layers
  int Nxt, Nyt, Nzt;        // with ghosts
  int64_t Nxyt, Ntotal;
  double ds;
  std::vector<uint8_t> lab;         // Ntotal labels
  std::vector<uint32_t> code;       // Ntotal; fluid cells only are meaningful
  // walls in the reference's vecWall order
  std::vector<int32_t> wall_cell;
  std::vector<uint8_t> wall_label;
  std::vector<double> wall_e1, wall_e2;
  std::vector<int32_t> wall_corner; // 3-fold corner id (0-based) or -1
  std::vector<uint64_t> wall_views; // 8 octants x (y code, z code) nibbles
  std::vector<int32_t> wid;         // Ntotal: wall index or -1
  int ncorner = 0;
  // Walls whose extrapolation reads another wall: the reference's serial wall
  // loop gives them the new values of earlier walls and the old values of
  // later ones.  They run after the others in dependency levels (a wall's
  // level is one more than the highest level of the earlier walls it reads);
  // reads of later walls come from a snapshot taken before the wall pass.
  std::vector<int32_t> coupled;           // wall indices, sorted by level
  std::vector<int32_t> coupled_level_off; // offsets of each level in `coupled`
  std::vector<int32_t> coupled_snap;      // 6 per coupled wall: snapshot index or -1
  std::vector<int32_t> snap_cell;         // cells to snapshot before the wall pass
  std::vector<int32_t> uncoupled;         // wall indices processed in parallel, grouped
                                          // by face count (1, 2, 3), raster order inside
  int32_t uncoupled_off[4] = {0, 0, 0, 0};  // group of walls with n faces: [off[n-1], off[n])
  // per wall, kWallPos storage positions: the wall cell in families 0..3,
  // then per face (up to 3), per distance (1, 2), per incoming octant q (the
  // face's in[q]) the neighbour cell in that octant's family
  static constexpr int kWallPos = 4 + 3 * 2 * 4;
  std::vector<int32_t> wall_pos;
  // fluid cells
  std::vector<int32_t> fluid_interior;    // not on any boundary plane: moments from sweep partials
  std::vector<int32_t> fluid_boundary;    // on i=1,Nx / j=1,Ny / k=1,Nz planes
  std::vector<int32_t> commit_walls;      // walls with a switch in some octant
  std::vector<int32_t> surface;           // cells on the surface of [1,Nx] x [1,Ny] x [1,Nz]
  int64_t nfluid = 0;

  // Storage of f.  Per octant only the cells that some kernel writes are
  // kept: fluid and wall cells, the x ghost planes 0 and Nx+1, and the
  // boundary planes x = 1, Nx, y = 1, Ny, z = 1, Nz.  Every other cell
  // (interior solid, y/z ghosts, x ghosts -1 and Nx+2) is zero throughout and
  // reads from one extra zero slot at index nstore.  Octant o keeps its cells
  // in the order of its family ofam(o): ascending sx*i + sy*j + sz*k with the
  // family's signs, then k, then j, so that one hyperplane of a sweep is one
  // contiguous range.  pos[F] maps a cell to its index in family F's order.
  int64_t nstore = 0;
  std::vector<int32_t> pos[4];

  // Sweep schedule.  Octant o updates its swept cells (fluid cells past the
  // octant's inflow layer) in order of the oriented hyperplane
  // s = oi + oj + ok, where oi = i-1 when the octant's cx > 0 and Nx-i when
  // cx < 0 (likewise oj, ok): a cell's upwind neighbours lie on s-1 and s-2,
  // so one hyperplane of all eight octants can be updated at once.
  // hp_task[8t .. 8t+8) describes task t: [0] storage index | octant << 28,
  // [1..6] storage indices of the upwind neighbours at distance 1 and 2 along
  // x, y, z, [7] cell | flags << 28, flags = the octant's first-order flags
  // for x, y, z (bits 0-2) and its "reads a switched wall" flag (bit 3).
  // Tasks of hyperplane s are hp_off[s] .. hp_off[s+1), octant-major, each
  // octant's tasks in ascending storage order.
  static constexpr int kTaskWords = 8;
  int nhp = 0;
  std::vector<int32_t> hp_task;
  std::vector<int64_t> hp_off;
  // hp_oct[8s + o]: first task of octant o in hyperplane s (hp_oct[8 nhp] =
  // the task count), so that a subset of the octants can be swept alone
  std::vector<int64_t> hp_oct;
  // Storage window of each (hyperplane, octant): every storage index its
  // tasks touch (the zero slot aside) lies in [hp_lo, hp_lo + hp_len).
  std::vector<int32_t> hp_lo, hp_len;  // nhp * 8

  // Tiled sweep schedule (build_tile_schedule).  Octant o's oriented grid
  // (oi, oj, ok in [0, N-1], as above) is cut into tiles of tx * ty * tz
  // cells; tile (I, J, K) reads upwind values only from itself and from tiles
  // (I-1, J, K), (I, J-1, K), (I, J, K-1), so the tiles of one diagonal
  // I + J + K (all octants) are independent and the diagonals run in order.
  // One wave walks a tile's planes ok = K*tz .. in order, the tx * ty cells of
  // a plane in raster order (oi fastest).  Each tile has (tz + 1) records of
  // kTileRecDw dwords: a header (octant, plane count, storage window) with the
  // two planes below the tile, then one record per plane (see kTileRec* for
  // the dword positions).  Storage positions are relative to the window's
  // first index; the zero slot is encoded as the window length.  Only tiles
  // with at least one swept cell are listed, ordered by diagonal, octant, K,
  // J, I; tl_off[D] is the first tile of diagonal D.
  static constexpr int kTileRecDw = 64;
  int tl_tx = 0, tl_ty = 0, tl_tz = 0, tl_ndiag = 0;
  int64_t tl_ntile = 0;
  std::vector<uint32_t> tl_rec;
  std::vector<int64_t> tl_off;
  int64_t tl_maxwin = 0;               // largest storage window (cells)
  int64_t tl_maxspan = 0;              // largest span of a tile's swept cell indices

  int64_t cell(int i, int j, int k) const {
    return (int64_t)(i + 1) + (int64_t)(j + 1) * Nxt + (int64_t)(k + 1) * Nxyt;
  }
};

// Returns 0 on success; on failure returns a PIKS_E* code and fills err.
int build_geometry(const GeomInput& in, const uint8_t* raw, int64_t nbytes, Geometry& g,
                   std::string& err);

// Fill the storage order (nstore, pos) and the sweep schedule (hp_task, hp_off).
void build_sweep_schedule(Geometry& g, int threads);

// Dword positions inside the tiled sweep's records (tile of tx * ty cells per
// plane, ns = tx * ty slots, slot s = b * tx + a for oriented offsets a, b).
// Header record: [0] octant | planes << 8, [1] window's first storage index,
// [2] window length, [3] smallest cell index of a swept cell, [4] the span of
// the swept cells' indices, [8 + s] plane ok = K*tz - 2, [8 + ns + s] plane
// K*tz - 1.
// Plane record: [0] mask of the slots this octant sweeps, [1 + s] the slot's
// storage position, [1 + ns + 2b + (a' + 2)] x halo a' = -2, -1 of row b,
// [1 + ns + 2ty + 2a + (b' + 2)] y halo b' = -2, -1 of column a,
// [1 + 2ns + 2ty + 2tx + s] cell | flags << 28 of a swept slot (flags as in
// hp_task).
constexpr int tile_rec_zhalo(int ns, int which) { return 8 + which * ns; }
constexpr int tile_rec_slot(int s) { return 1 + s; }
constexpr int tile_rec_xhalo(int ns, int q) { return 1 + ns + q; }
constexpr int tile_rec_yhalo(int ns, int ty, int q) { return 1 + ns + 2 * ty + q; }
constexpr int tile_rec_cell(int ns, int tx, int ty, int s) { return 1 + ns + 2 * ty + 2 * tx + s; }

// The hyperplane sweep's task records as its kernels read them (in place):
// element offsets inside the (hyperplane, octant) storage window, with nc8
// velocities per octant, the zero slot as the window's end.
void make_window_relative(Geometry& g, int nc8, int threads);

// Records of the single-launch sweep, from the window-relative hyperplane
// records: the tasks in octant-major, hyperplane order, kTaskWords words
// each (off[o] the first task of octant o, off[8] the count).  With
// strips > 1 an octant's tasks go strip by strip (strips equal bands of y,
// in the octant's y direction), in hyperplane order within a strip: every
// task still follows its upwind neighbours, and the tasks written and not
// yet read at any time span two hyperplanes of one strip, not of the image.  Word 0 is the
// task's storage index; words 1..6 its upwind sources (flow_switched_word's
// encoding; bit 31 with bits 28..30 clear marks the nearest upwind
// neighbour along each axis that the octant sweeps, whose flag the task
// waits for); word 7 the cell and flags, bit 31 cleared once the task's
// reads of switched wall slots are resolved into words 3..6.  Returns 0, or
// PIKS_E_GEOM with err filled.
int build_flow_records(const Geometry& g, int nc8, std::vector<int32_t>& rec, int64_t off[9], std::string& err,
                       int strips = 1);

// Source word of a single-launch sweep record for the read of neighbour n
// (a y neighbour for axis 1, z for axis 2) by a task of octant o, when the
// reference switched that wall slot (the views of Geometry::wall_views, read
// as the sweep's switched_source reads them): bits 0..27 an index, 28..30 x,
// bit 31 c; c = 0: storage index of octant o ^ x (the switched octant's own
// slot of n), c = 1 and x = 1: corner store entry corner * 8 + o.  Returns
// -1 when the plain slot applies, -2 for a corner view without a corner.
int64_t flow_switched_word(const Geometry& g, int o, int64_t n, int axis);

// The stored slots (octant o, storage index p) that octant o's sweep does not
// update (walls, ghost planes, inflow layers, ...), as p | o << 28 sorted by
// octant and index: the slots a defect-correction iteration updates outside
// the sweep.  Needs the storage order.
std::vector<uint32_t> build_nonswept(const Geometry& g, int threads);

// Fill the tiled sweep schedule (tl_*); needs the storage order.  Returns
// false when a tile shape does not fit the record (the schedule stays empty).
bool build_tile_schedule(Geometry& g, int tx, int ty, int tz, int threads);

template <typename F>
void parallel_for(int n, int threads, F&& f) {
  threads = std::max(1, std::min(threads, n));
  if (threads == 1) {
    for (int q = 0; q < n; ++q) f(q);
    return;
  }
  std::atomic<int> next{0};
  auto run = [&] {
    for (int q; (q = next.fetch_add(1)) < n;) f(q);
  };
  std::vector<std::thread> pool;
  pool.reserve((size_t)threads - 1);
  for (int t = 1; t < threads; ++t) pool.emplace_back(run);
  run();
  for (auto& t : pool) t.join();
}

}  // namespace piks
