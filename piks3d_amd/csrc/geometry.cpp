// geometry.cpp -- image -> device layouts (see geometry.h).
#include "geometry.h"

#include <algorithm>
#include <climits>
#include <array>
#include <cstdlib>
#include <cstring>
#include <unordered_map>

#include "../../include/piks3d.h"
#include "wallprog.h"

namespace piks {

namespace {
thread_local int tl_host_threads = 0;
}

int host_threads() {
  if (tl_host_threads > 0) return tl_host_threads;
  if (const char* e = std::getenv("PIKS_HOST_THREADS")) {
    const int n = std::atoi(e);
    if (n > 0) return std::min(n, 64);
  }
  const unsigned hw = std::thread::hardware_concurrency();
  return (int)std::max(1u, std::min(hw, 16u));
}

void set_host_threads(int n) { tl_host_threads = n > 0 ? n : 0; }

namespace {

struct Grid {
  int Nx, Ny, Nz, Nxt;
  int64_t Nxyt;
  std::vector<int8_t> a;
  int8_t& at(int i, int j, int k) {
    return a[(size_t)((i + 1) + (int64_t)(j + 1) * Nxt + (int64_t)(k + 1) * Nxyt)];
  }
};

bool outside(const Grid& g, int i, int j, int k) {
  return i < 1 || i > g.Nx || j < 1 || j > g.Ny || k < 1 || k > g.Nz;
}

void set_ghosts(Grid& g, int8_t v) {
  for (int k = -1; k <= g.Nz + 2; ++k)
    for (int j = -1; j <= g.Ny + 2; ++j)
      for (int i = -1; i <= g.Nx + 2; ++i)
        if (outside(g, i, j, k)) g.at(i, j, k) = v;
}

// One in-place drilling pass over i = 2..Nx-1 (src/physicalGrid.F90:352-377):
// a fluid voxel closed on both sides along some axis opens a 2x2x2 block,
// shifted back at the j = Ny and k = Nz faces.
void drill(Grid& g) {
  for (int k = 1; k <= g.Nz; ++k)
    for (int j = 1; j <= g.Ny; ++j)
      for (int i = 2; i <= g.Nx - 1; ++i) {
        if (g.at(i, j, k) != FLUID) continue;
        bool closed = (g.at(i + 1, j, k) != FLUID && g.at(i - 1, j, k) != FLUID) ||
                      (g.at(i, j + 1, k) != FLUID && g.at(i, j - 1, k) != FLUID) ||
                      (g.at(i, j, k + 1) != FLUID && g.at(i, j, k - 1) != FLUID);
        if (!closed) continue;
        int j0 = j == g.Ny ? j - 1 : j;
        int k0 = k == g.Nz ? k - 1 : k;
        for (int kk = k0; kk <= k0 + 1; ++kk)
          for (int jj = j0; jj <= j0 + 1; ++jj)
            for (int ii = i; ii <= i + 1; ++ii) g.at(ii, jj, kk) = FLUID;
      }
}

// One in-place pass turning one-voxel-thick solid into fluid, with the ghost
// ring counted as fluid (src/physicalGrid.F90:382-422).
void thin_walls(Grid& g) {
  set_ghosts(g, FLUID);
  for (int k = 1; k <= g.Nz; ++k)
    for (int j = 1; j <= g.Ny; ++j)
      for (int i = 1; i <= g.Nx; ++i) {
        if (g.at(i, j, k) != SOLID) continue;
        if ((g.at(i + 1, j, k) == FLUID && g.at(i - 1, j, k) == FLUID) ||
            (g.at(i, j + 1, k) == FLUID && g.at(i, j - 1, k) == FLUID) ||
            (g.at(i, j, k + 1) == FLUID && g.at(i, j, k - 1) == FLUID))
          g.at(i, j, k) = FLUID;
      }
  set_ghosts(g, GHOST);
}

constexpr int kSideE = 1, kSideW = 2, kSideN = 4, kSideS = 8, kSideF = 16, kSideB = 32;

}  // namespace

int build_geometry(const GeomInput& in, const uint8_t* raw, int64_t nbytes, Geometry& g,
                   std::string& err) {
  if (in.nx < 1 || in.ny < 3 || in.nz < 3 || in.fluid_layer < 0) {
    err = "domain too small: need Nx >= 1, Ny >= 3, Nz >= 3";
    return PIKS_E_ARG;
  }
  const int threads = in.threads > 0 ? in.threads : host_threads();
  const int fl = in.fluid_layer;
  g.Nx = in.nx + 2 * fl;
  g.Ny = in.ny;
  g.Nz = in.nz;
  if (g.Nx < 4) {
    err = "domain too small: need Nx + 2*fluidLayer >= 4";
    return PIKS_E_ARG;
  }
  g.Nxt = g.Nx + 4;
  g.Nyt = g.Ny + 4;
  g.Nzt = g.Nz + 4;
  g.Nxyt = (int64_t)g.Nxt * g.Nyt;
  g.Ntotal = g.Nxyt * g.Nzt;
  if (g.Ntotal >= (int64_t)1 << 31) {
    err = "grid too large for 32-bit cell indices";
    return PIKS_E_ARG;
  }
  g.ds = 1.0 / (g.Nx - 1 - 2 * fl);  // src/physicalGrid.F90:228

  const int64_t need = (int64_t)in.nx * in.ny * in.nz;
  if (nbytes < need) {
    err = "image has " + std::to_string(nbytes) + " bytes, need " + std::to_string(need);
    return PIKS_E_IMAGE;
  }
  Grid G{g.Nx, g.Ny, g.Nz, g.Nxt, g.Nxyt, std::vector<int8_t>((size_t)g.Ntotal, FLUID)};
  int64_t pos = 0;
  for (int k = 1; k <= g.Nz; ++k)
    for (int j = 1; j <= g.Ny; ++j)
      for (int i = fl + 1; i <= g.Nx - fl; ++i) {
        uint8_t v = raw[pos++];
        if (v > 1) {
          err = "invalid image data file: byte " + std::to_string((int)v) + " at offset " +
                std::to_string(pos - 1);
          return PIKS_E_IMAGE;
        }
        G.at(i, j, k) = v ? SOLID : FLUID;
      }
  set_ghosts(G, GHOST);
  drill(G);
  thin_walls(G);
  drill(G);
  thin_walls(G);

  // labels (src/physicalGrid.F90:511-671)
  g.lab.assign((size_t)g.Ntotal, GHOST);
  parallel_for(g.Nz, threads, [&](int kk) {
    const int k = kk + 1;
    for (int j = 1; j <= g.Ny; ++j)
      for (int i = 1; i <= g.Nx; ++i) {
        int8_t v = G.at(i, j, k);
        int lab = v;
        if (v == SOLID) {
          int p = 1;
          if (G.at(i + 1, j, k) == FLUID) p *= 2;
          if (G.at(i - 1, j, k) == FLUID) p *= 3;
          if (G.at(i, j + 1, k) == FLUID) p *= 5;
          if (G.at(i, j - 1, k) == FLUID) p *= 7;
          if (G.at(i, j, k + 1) == FLUID) p *= 9;
          if (G.at(i, j, k - 1) == FLUID) p *= 11;
          int l = label_of_product(p);
          if (l > 0) lab = l;
        }
        g.lab[(size_t)g.cell(i, j, k)] = (uint8_t)lab;
      }
  });

  // walls in raster order, 3-fold corner ids (src/physicalGrid.F90:701-748)
  g.wid.assign((size_t)g.Ntotal, -1);
  int near_x = 0;
  for (int k = 1; k <= g.Nz; ++k)
    for (int j = 1; j <= g.Ny; ++j)
      for (int i = 1; i <= g.Nx; ++i) {
        int64_t c = g.cell(i, j, k);
        int lab = g.lab[(size_t)c];
        if (!is_wall(lab)) continue;
        g.wid[(size_t)c] = (int32_t)g.wall_cell.size();
        g.wall_cell.push_back((int32_t)c);
        g.wall_label.push_back((uint8_t)lab);
        if (is_corner3(lab)) {
          g.wall_corner.push_back(g.ncorner++);
          if (i <= 2 || i >= g.Nx - 1) ++near_x;
        } else {
          g.wall_corner.push_back(-1);
        }
      }
  if (near_x > 0) {
    // the reference's x self-message would carry these corners while the
    // receive does not (src/flow.F90:124-127 vs src/physicalGrid.F90:835-885):
    // MPI aborts the reference on such an image
    err = "3-fold corner within two layers of the x boundary (the reference aborts in MPI): "
          "add fluid layers";
    return PIKS_E_GEOM;
  }
  const size_t nwall = g.wall_cell.size();

  // extrapolation coefficients (src/physicalGrid.F90:891-948).  The order-3
  // rule decodes each wall's position from its 1-based local id with the x
  // index landing one cell east; that decode is reproduced so results match.
  g.wall_e1.assign(nwall, 2.0);
  g.wall_e2.assign(nwall, -1.0);
  if (in.wall_ext_order == 1) {
    g.wall_e1.assign(nwall, 1.0);
    g.wall_e2.assign(nwall, 0.0);
  } else if (in.wall_ext_order != 2) {
    const int Nxt = g.Nxt;
    const int Nxy = (int)g.Nxyt;
    for (size_t w = 0; w < nwall; ++w) {
      int localid = g.wall_cell[w] + 1;
      int k = -1 + localid / Nxy;
      int j = -1 + (localid - (k + 1) * Nxy) / Nxt;
      int i = -1 + localid - (k + 1) * Nxy - (j + 1) * Nxt;
      int sides = wall_sides(g.lab[(size_t)g.cell(i, j, k)]);
      bool first = ((sides & kSideW) && i == 2 && g.lab[(size_t)g.cell(1, j, k)] == FLUID) ||
                   ((sides & kSideE) && i == g.Nx - 1 && g.lab[(size_t)g.cell(g.Nx, j, k)] == FLUID) ||
                   ((sides & kSideS) && j == 2 && g.lab[(size_t)g.cell(i, 1, k)] == FLUID) ||
                   ((sides & kSideN) && j == g.Ny - 1 && g.lab[(size_t)g.cell(i, g.Ny, k)] == FLUID) ||
                   ((sides & kSideB) && k == 2 && g.lab[(size_t)g.cell(i, j, 1)] == FLUID) ||
                   ((sides & kSideF) && k == g.Nz - 1 && g.lab[(size_t)g.cell(i, j, g.Nz)] == FLUID);
      if (first) {
        g.wall_e1[w] = 1.0;
        g.wall_e2[w] = 0.0;
      }
    }
  }

  // sweep membership: fluid, interior, and not on the first layer of the
  // octant's upwind side (src/physicalGrid.F90:977-985 and analogues)
  auto in_list = [&](int o, int64_t c) -> bool {
    if (g.lab[(size_t)c] != FLUID) return false;
    int k = (int)(c / g.Nxyt) - 1;
    int j = (int)((c % g.Nxyt) / g.Nxt) - 1;
    int i = (int)(c % g.Nxt) - 1;
    if (i < 1 || i > g.Nx || j < 1 || j > g.Ny || k < 1 || k > g.Nz) return false;
    if (osx(o) > 0 ? i < 2 : i > g.Nx - 1) return false;
    if (osy(o) > 0 ? j < 2 : j > g.Ny - 1) return false;
    if (osz(o) > 0 ? k < 2 : k > g.Nz - 1) return false;
    return true;
  };

  // per wall and octant: which value a fluid cell reading it along y or z
  // sees.  The reference switches the wall's slot in place when the sweep
  // reaches the cell just downstream in y or z (src/solver.F90:184-204 ...);
  // raster order makes x readers see the unswitched slot, y readers the
  // y-switched one, z readers the z-switched one, else the y-switched one.
  g.wall_views.assign(nwall, 0);
  std::vector<uint8_t> commits(nwall, 0);
  const int kWallChunk = 4096;
  parallel_for((int)((nwall + kWallChunk - 1) / kWallChunk), threads, [&](int chunk) {
  for (size_t w = (size_t)chunk * kWallChunk; w < std::min(nwall, (size_t)(chunk + 1) * kWallChunk); ++w) {
    int64_t c = g.wall_cell[w];
    int lab = g.wall_label[w];
    uint64_t v = 0;
    bool any = false;
    for (int o = 0; o < 8; ++o) {
      int ys = switch_source(o, 1, lab), zs = switch_source(o, 2, lab);
      bool ytrig = ys != 0 && in_list(o, c + (int64_t)osy(o) * g.Nxt);
      bool ztrig = zs != 0 && in_list(o, c + (int64_t)osz(o) * g.Nxyt);
      uint64_t yc = ytrig ? (uint64_t)ys : 0;
      uint64_t zc = ztrig ? (uint64_t)zs : yc;
      v |= (yc << (8 * o)) | (zc << (8 * o + 4));
      any = any || zc != 0;
    }
    g.wall_views[w] = v;
    commits[w] = any;
  }
  });
  for (size_t w = 0; w < nwall; ++w)
    if (commits[w]) g.commit_walls.push_back((int32_t)w);

  // per fluid cell: first-order flags and the switched-wall flag per octant
  // (src/physicalGrid.F90:1254-1622: first order along an axis when the cell
  // two upstream is a ghost or the upstream cell is a wall facing this one)
  g.code.assign((size_t)g.Ntotal, 0);
  const int face_toward[3][2] = {{kSideE, kSideW}, {kSideN, kSideS}, {kSideF, kSideB}};
  const int64_t stride[3] = {1, g.Nxt, g.Nxyt};
  // one z plane per work item; the fluid lists are joined in plane order, so
  // they keep the raster order of a serial pass
  std::vector<std::vector<int32_t>> plane_int((size_t)g.Nz + 1), plane_bnd((size_t)g.Nz + 1);
  parallel_for(g.Nz, threads, [&](int kk) {
    const int k = kk + 1;
    std::vector<int32_t>& l_int = plane_int[(size_t)k];
    std::vector<int32_t>& l_bnd = plane_bnd[(size_t)k];
    for (int j = 1; j <= g.Ny; ++j)
      for (int i = 1; i <= g.Nx; ++i) {
        int64_t c = g.cell(i, j, k);
        if (g.lab[(size_t)c] != FLUID) continue;
        bool bnd = i == 1 || i == g.Nx || j == 1 || j == g.Ny || k == 1 || k == g.Nz;
        (bnd ? l_bnd : l_int).push_back((int32_t)c);
        uint32_t code = 0;
        for (int o = 0; o < 8; ++o) {
          int s[3] = {osx(o), osy(o), osz(o)};
          bool first[3];
          for (int a = 0; a < 3; ++a) {
            int64_t up1 = c - s[a] * stride[a], up2 = c - 2 * s[a] * stride[a];
            first[a] = g.lab[(size_t)up2] == GHOST ||
                       (wall_sides(g.lab[(size_t)up1]) & face_toward[a][s[a] > 0 ? 0 : 1]) != 0;
            if (first[a]) code |= 1u << (3 * o + a);
          }
          if (!in_list(o, c)) continue;
          bool view = false;
          for (int a = 1; a <= 2; ++a)
            for (int d = 1; d <= 2; ++d) {
              if (d == 2 && first[a]) continue;  // zero weight
              int64_t n = c - d * s[a] * stride[a];
              int w = g.wid[(size_t)n];
              if (w < 0) continue;
              if ((g.wall_views[w] >> (8 * o + (a == 2 ? 4 : 0))) & 15u) view = true;
            }
          if (view) code |= 1u << (24 + o);
        }
        g.code[(size_t)c] = code;
      }
  });
  for (int k = 1; k <= g.Nz; ++k) {
    g.fluid_interior.insert(g.fluid_interior.end(), plane_int[(size_t)k].begin(), plane_int[(size_t)k].end());
    g.fluid_boundary.insert(g.fluid_boundary.end(), plane_bnd[(size_t)k].begin(), plane_bnd[(size_t)k].end());
  }
  g.nfluid = (int64_t)(g.fluid_interior.size() + g.fluid_boundary.size());
  for (int k = 1; k <= g.Nz; ++k)
    for (int j = 1; j <= g.Ny; ++j)
      for (int i = 1; i <= g.Nx; ++i)
        if (i == 1 || i == g.Nx || j == 1 || j == g.Ny || k == 1 || k == g.Nz)
          g.surface.push_back((int32_t)g.cell(i, j, k));

  // walls whose extrapolation reads another wall (see geometry.h)
  std::vector<int> level(nwall, 0);
  std::vector<std::array<int32_t, 6>> slots_of(nwall);
  std::unordered_map<int32_t, int32_t> snap_of_cell;
  std::vector<char> is_coupled(nwall, 0);
  int maxlevel = 0;
  for (size_t w = 0; w < nwall; ++w) {
    WallProg P = wall_prog(g.wall_label[w]);
    int64_t c = g.wall_cell[w];
    std::array<int32_t, 6> slots = {-1, -1, -1, -1, -1, -1};
    int lev = 0;
    for (int a = 0; a < P.nface; ++a)
      for (int d = 1; d <= 2; ++d) {
        int64_t n = c + (int64_t)d * P.face[a].dir * stride[(int)P.face[a].axis];
        int wn = g.wid[(size_t)n];
        if (wn < 0) continue;
        is_coupled[w] = 1;
        if ((size_t)wn > w) {  // later in the wall order: needs its old value
          auto it = snap_of_cell.find((int32_t)n);
          int32_t idx;
          if (it == snap_of_cell.end()) {
            idx = (int32_t)g.snap_cell.size();
            snap_of_cell[(int32_t)n] = idx;
            g.snap_cell.push_back((int32_t)n);
          } else {
            idx = it->second;
          }
          slots[a * 2 + (d - 1)] = idx;
        } else {
          lev = std::max(lev, level[wn]);  // earlier wall: must run before
        }
      }
    if (is_coupled[w]) {
      level[w] = lev + 1;
      maxlevel = std::max(maxlevel, level[w]);
      slots_of[w] = slots;
    } else {
      g.uncoupled.push_back((int32_t)w);
    }
  }
  // the uncoupled walls are independent of each other: group them by face
  // count so that each group runs in a kernel sized for its register need
  {
    std::vector<int32_t> by_nf[3];
    for (int32_t w : g.uncoupled) by_nf[wall_prog(g.wall_label[(size_t)w]).nface - 1].push_back(w);
    g.uncoupled.clear();
    for (int q = 0; q < 3; ++q) {
      g.uncoupled_off[q] = (int32_t)g.uncoupled.size();
      g.uncoupled.insert(g.uncoupled.end(), by_nf[q].begin(), by_nf[q].end());
    }
    g.uncoupled_off[3] = (int32_t)g.uncoupled.size();
  }
  g.coupled_level_off.push_back(0);
  for (int L = 1; L <= maxlevel; ++L) {
    for (size_t w = 0; w < nwall; ++w)
      if (is_coupled[w] && level[w] == L) {
        g.coupled.push_back((int32_t)w);
        for (int q = 0; q < 6; ++q) g.coupled_snap.push_back(slots_of[w][q]);
      }
    g.coupled_level_off.push_back((int32_t)g.coupled.size());
  }
  if (g.Ntotal >= ((int64_t)1 << 28)) {
    err = "domain of " + std::to_string(g.Ntotal) + " cells (with ghosts) is too large";
    return PIKS_E_ARG;
  }
  build_sweep_schedule(g, threads);
  return PIKS_OK;
}

void build_wall_positions(Geometry& g, int threads);

void build_sweep_schedule(Geometry& g, int threads) {
  // cells that are stored
  std::vector<uint8_t> stored((size_t)g.Ntotal, 0);
  for (int k = 1; k <= g.Nz; ++k)
    for (int j = 1; j <= g.Ny; ++j)
      for (int i = 0; i <= g.Nx + 1; ++i) {
        const uint8_t l = g.lab[(size_t)g.cell(i, j, k)];
        const bool plane = i <= 1 || i >= g.Nx || j == 1 || j == g.Ny || k == 1 || k == g.Nz;
        if (plane || l == FLUID || l >= 20) stored[(size_t)g.cell(i, j, k)] = 1;
      }
  int64_t ns = 0;
  for (uint8_t v : stored) ns += v;
  g.nstore = ns;
  // family orders: counting sort on sx*i + sy*j + sz*k, stable over (k, j)
  const int kmin = -(g.Nx + 2) - (g.Ny + 2) - (g.Nz + 2), nkey = 2 * (-kmin) + 1;
  parallel_for(4, threads, [&](int F) {
    const int sx = osx(F), sy = osy(F), sz = osz(F);
    std::vector<int64_t> start((size_t)nkey + 1, 0);
    for (int k = -1; k <= g.Nz + 2; ++k)
      for (int j = -1; j <= g.Ny + 2; ++j)
        for (int i = -1; i <= g.Nx + 2; ++i)
          if (stored[(size_t)g.cell(i, j, k)]) ++start[(size_t)(sx * i + sy * j + sz * k - kmin) + 1];
    for (int q = 0; q < nkey; ++q) start[(size_t)q + 1] += start[(size_t)q];
    std::vector<int32_t>& pos = g.pos[F];
    pos.assign((size_t)g.Ntotal, (int32_t)ns);
    for (int k = -1; k <= g.Nz + 2; ++k)
      for (int j = -1; j <= g.Ny + 2; ++j)
        for (int i = -1; i <= g.Nx + 2; ++i) {
          const int64_t c = g.cell(i, j, k);
          if (stored[(size_t)c]) pos[(size_t)c] = (int32_t)start[(size_t)(sx * i + sy * j + sz * k - kmin)]++;
        }
  });

  // tasks, bucketed by (hyperplane, octant); raster order (k, j) inside a
  // bucket is ascending storage order
  g.nhp = (g.Nx - 1) + (g.Ny - 1) + (g.Nz - 1) + 1;
  const int64_t nb = (int64_t)g.nhp * 8;
  std::vector<int64_t> count((size_t)nb + 1, 0);
  auto oriented = [&](int o, int i, int j, int k, int* s) {
    const int oi = osx(o) > 0 ? i - 1 : g.Nx - i;
    const int oj = osy(o) > 0 ? j - 1 : g.Ny - j;
    const int ok = osz(o) > 0 ? k - 1 : g.Nz - k;
    *s = oi + oj + ok;
    return oi >= 1 && oj >= 1 && ok >= 1;  // the octant's inflow layer is not swept
  };
  const int W = Geometry::kTaskWords;
  for (int pass = 0; pass < 2; ++pass) {
    std::vector<int64_t> at;
    if (pass == 1) {
      for (int64_t b = 0; b < nb; ++b) count[(size_t)b + 1] += count[(size_t)b];
      g.hp_off.assign((size_t)g.nhp + 1, 0);
      for (int s = 0; s <= g.nhp; ++s) g.hp_off[(size_t)s] = count[(size_t)s * 8];
      g.hp_task.assign((size_t)count[(size_t)nb] * W, 0);
      at.assign(count.begin(), count.end() - 1);
    }
    // the octants fill disjoint buckets (b % 8 == o): one work item each
    parallel_for(8, threads, [&](int o) {
      const std::vector<int32_t>& pos = g.pos[ofam(o)];
      const int64_t dx = osx(o), dy = (int64_t)osy(o) * g.Nxt, dz = (int64_t)osz(o) * g.Nxyt;
      for (int k = 1; k <= g.Nz; ++k)
        for (int j = 1; j <= g.Ny; ++j)
          for (int i = 1; i <= g.Nx; ++i) {
            const int64_t c = g.cell(i, j, k);
            int s;
            if (g.lab[(size_t)c] != FLUID || !oriented(o, i, j, k, &s)) continue;
            const int64_t b = (int64_t)s * 8 + o;
            if (pass == 0) {
              ++count[(size_t)b + 1];
              continue;
            }
            int32_t* t = &g.hp_task[(size_t)(at[(size_t)b]++ * W)];
            const uint32_t cc = g.code[(size_t)c];
            const uint32_t flags = ((cc >> (3 * o)) & 7u) | (((cc >> (24 + o)) & 1u) << 3);
            t[0] = (int32_t)((uint32_t)pos[(size_t)c] | (uint32_t)o << 28);
            t[1] = pos[(size_t)(c - dx)];
            t[2] = pos[(size_t)(c - 2 * dx)];
            t[3] = pos[(size_t)(c - dy)];
            t[4] = pos[(size_t)(c - 2 * dy)];
            t[5] = pos[(size_t)(c - dz)];
            t[6] = pos[(size_t)(c - 2 * dz)];
            t[7] = (int32_t)((uint32_t)c | flags << 28);
          }
    });
  }
  g.hp_oct.assign(count.begin(), count.end());
  // storage windows
  g.hp_lo.assign((size_t)nb, 0);
  g.hp_len.assign((size_t)nb, 0);
  const int kChunk = 64;
  parallel_for((int)((nb + kChunk - 1) / kChunk), threads, [&](int chunk) {
  for (int64_t b = (int64_t)chunk * kChunk; b < std::min<int64_t>(nb, (int64_t)(chunk + 1) * kChunk); ++b) {
    int32_t lo = INT32_MAX, hi = -1;
    for (int64_t t = count[(size_t)b]; t < count[(size_t)b + 1]; ++t) {
      const int32_t* r = &g.hp_task[(size_t)(t * W)];
      for (int q = 0; q < 7; ++q) {
        const int32_t v = q == 0 ? (int32_t)((uint32_t)r[0] & 0x0fffffffu) : r[q];
        if (v == (int32_t)ns) continue;
        lo = std::min(lo, v);
        hi = std::max(hi, v);
      }
    }
    if (hi >= lo) {
      g.hp_lo[(size_t)b] = lo;
      g.hp_len[(size_t)b] = hi - lo + 1;
    }
  }
  });

  build_wall_positions(g, threads);
}

bool build_tile_schedule(Geometry& g, int TX, int TY, int TZ, int threads) {
  const int NS = TX * TY;
  const int W = Geometry::kTileRecDw;
  g.tl_rec.clear();
  g.tl_off.clear();
  g.tl_ntile = 0;
  g.tl_ndiag = 0;
  g.tl_maxwin = 0;
  if (TX < 2 || TY < 2 || TZ < 2 || tile_rec_cell(NS, TX, TY, NS) > W || tile_rec_zhalo(NS, 2) > W ||
      NS > 32)
    return false;
  g.tl_tx = TX; g.tl_ty = TY; g.tl_tz = TZ;
  const int NTX = (g.Nx + TX - 1) / TX, NTY = (g.Ny + TY - 1) / TY, NTZ = (g.Nz + TZ - 1) / TZ;
  const int ND = NTX + NTY + NTZ - 2;
  const int64_t zero = g.nstore;
  // physical cell of oriented (oi, oj, ok) in octant o; the oriented range
  // [-2, N+1] covers the ghosted grid
  auto phys = [&](int o, int oi, int oj, int ok) {
    const int i = osx(o) > 0 ? oi + 1 : g.Nx - oi;
    const int j = osy(o) > 0 ? oj + 1 : g.Ny - oj;
    const int k = osz(o) > 0 ? ok + 1 : g.Nz - ok;
    return g.cell(i, j, k);
  };
  auto inside = [&](int oi, int oj, int ok) {
    return oi >= 0 && oi < g.Nx && oj >= 0 && oj < g.Ny && ok >= 0 && ok < g.Nz;
  };
  auto swept = [&](int o, int oi, int oj, int ok) {
    return inside(oi, oj, ok) && oi >= 1 && oj >= 1 && ok >= 1 && g.lab[(size_t)phys(o, oi, oj, ok)] == FLUID;
  };
  // tiles with a swept cell, per octant
  struct Tile { int32_t d, o, I, J, K; };
  std::vector<std::vector<Tile>> per(8);
  parallel_for(8, threads, [&](int o) {
    for (int K = 0; K < NTZ; ++K)
      for (int J = 0; J < NTY; ++J)
        for (int I = 0; I < NTX; ++I) {
          bool any = false;
          for (int c = 0; c < TZ && !any; ++c)
            for (int b = 0; b < TY && !any; ++b)
              for (int a = 0; a < TX && !any; ++a)
                any = swept(o, I * TX + a, J * TY + b, K * TZ + c);
          if (any) per[(size_t)o].push_back(Tile{I + J + K, o, I, J, K});
        }
  });
  std::vector<Tile> all;
  for (auto& v : per) all.insert(all.end(), v.begin(), v.end());
  // diagonal, then octant, then K, J, I (each octant's list is already in
  // K, J, I order, so a stable sort on (diagonal, octant) suffices)
  std::stable_sort(all.begin(), all.end(),
                   [](const Tile& x, const Tile& y) { return x.d != y.d ? x.d < y.d : x.o < y.o; });
  g.tl_ntile = (int64_t)all.size();
  g.tl_ndiag = ND;
  g.tl_off.assign((size_t)ND + 1, 0);
  for (const Tile& t : all) ++g.tl_off[(size_t)t.d + 1];
  for (int d = 0; d < ND; ++d) g.tl_off[(size_t)d + 1] += g.tl_off[(size_t)d];
  const size_t per_tile = (size_t)(TZ + 1) * W;
  g.tl_rec.assign(all.size() * per_tile, 0);
  std::atomic<int64_t> gmax{0}, gspan{0};
  const int kChunk = 256;
  parallel_for((int)((all.size() + kChunk - 1) / kChunk), threads, [&](int chunk) {
    int64_t local_max = 0, local_span = 0;
    std::vector<int64_t> posv;
    for (size_t t = (size_t)chunk * kChunk; t < std::min(all.size(), (size_t)(chunk + 1) * kChunk); ++t) {
      const Tile& T = all[t];
      const int o = T.o;
      const std::vector<int32_t>& pos = g.pos[ofam(o)];
      uint32_t* R = &g.tl_rec[t * per_tile];
      const int oi0 = T.I * TX, oj0 = T.J * TY, ok0 = T.K * TZ;
      const int np = std::min(TZ, g.Nz - ok0);
      // storage positions (absolute; the zero slot is `zero`), gathered first
      // to find the window
      auto sp = [&](int oi, int oj, int ok, bool need_inside) -> int64_t {
        if (need_inside && !inside(oi, oj, ok)) return zero;
        if (oi < -2 || oj < -2 || ok < -2 || oi > g.Nx + 1 || oj > g.Ny + 1 || ok > g.Nz + 1) return zero;
        return pos[(size_t)phys(o, oi, oj, ok)];
      };
      int64_t lo = INT64_MAX, hi = -1;
      auto note = [&](int64_t p) {
        if (p != zero) { lo = std::min(lo, p); hi = std::max(hi, p); }
        return p;
      };
      // header
      R[0] = (uint32_t)o | (uint32_t)np << 8;
      posv.clear();
      for (int w = 0; w < 2; ++w)
        for (int b = 0; b < TY; ++b)
          for (int a = 0; a < TX; ++a) {
            // planes below the tile: only cells inside the domain in x, y
            const int oi = oi0 + a, oj = oj0 + b, ok = ok0 - 2 + w;
            const bool ok_xy = oi < g.Nx && oj < g.Ny;
            posv.push_back(note(ok_xy ? sp(oi, oj, ok, false) : zero));
          }
      for (int c = 0; c < np; ++c) {
        const int ok = ok0 + c;
        for (int b = 0; b < TY; ++b)
          for (int a = 0; a < TX; ++a) posv.push_back(note(sp(oi0 + a, oj0 + b, ok, true)));
        for (int b = 0; b < TY; ++b)
          for (int ap = -2; ap <= -1; ++ap) {
            const int oj = oj0 + b;
            posv.push_back(note(oj < g.Ny ? sp(oi0 + ap, oj, ok, false) : zero));
          }
        for (int a = 0; a < TX; ++a)
          for (int bp = -2; bp <= -1; ++bp) {
            const int oi = oi0 + a;
            posv.push_back(note(oi < g.Nx ? sp(oi, oj0 + bp, ok, false) : zero));
          }
      }
      if (hi < lo) { lo = 0; hi = -1; }
      const int64_t len = hi - lo + 1;
      local_max = std::max(local_max, len);
      R[1] = (uint32_t)lo;
      R[2] = (uint32_t)len;
      // range of the swept cells' indices (the partial moments they write)
      int64_t clo = INT64_MAX, chi = -1;
      for (int c = 0; c < np; ++c)
        for (int b = 0; b < TY; ++b)
          for (int a = 0; a < TX; ++a)
            if (swept(o, oi0 + a, oj0 + b, ok0 + c)) {
              const int64_t cl = phys(o, oi0 + a, oj0 + b, ok0 + c);
              clo = std::min(clo, cl);
              chi = std::max(chi, cl);
            }
      R[3] = (uint32_t)clo;
      R[4] = (uint32_t)(chi - clo + 1);
      local_span = std::max(local_span, chi - clo + 1);
      auto rel = [&](int64_t p) { return (uint32_t)(p == zero ? len : p - lo); };
      size_t q = 0;
      for (int w = 0; w < 2; ++w)
        for (int s = 0; s < NS; ++s) R[tile_rec_zhalo(NS, w) + s] = rel(posv[q++]);
      for (int c = 0; c < np; ++c) {
        uint32_t* P = R + (size_t)(c + 1) * W;
        const int ok = ok0 + c;
        uint32_t mask = 0;
        for (int s = 0; s < NS; ++s) P[tile_rec_slot(s)] = rel(posv[q++]);
        for (int x = 0; x < 2 * TY; ++x) P[tile_rec_xhalo(NS, x)] = rel(posv[q++]);
        for (int y = 0; y < 2 * TX; ++y) P[tile_rec_yhalo(NS, TY, y)] = rel(posv[q++]);
        for (int b = 0; b < TY; ++b)
          for (int a = 0; a < TX; ++a) {
            const int s = b * TX + a;
            const int oi = oi0 + a, oj = oj0 + b;
            if (!swept(o, oi, oj, ok)) continue;
            mask |= 1u << s;
            const int64_t cl = phys(o, oi, oj, ok);
            const uint32_t cc = g.code[(size_t)cl];
            const uint32_t flags = ((cc >> (3 * o)) & 7u) | (((cc >> (24 + o)) & 1u) << 3);
            P[tile_rec_cell(NS, TX, TY, s)] = (uint32_t)cl | flags << 28;
          }
        P[0] = mask;
      }
    }
    int64_t cur = gmax.load();
    while (local_max > cur && !gmax.compare_exchange_weak(cur, local_max)) {
    }
    cur = gspan.load();
    while (local_span > cur && !gspan.compare_exchange_weak(cur, local_span)) {
    }
  });
  g.tl_maxwin = gmax.load();
  g.tl_maxspan = gspan.load();
  return true;
}

std::vector<uint32_t> build_nonswept(const Geometry& g, int threads) {
  std::vector<std::vector<uint32_t>> per(8);
  parallel_for(8, threads, [&](int o) {
    const std::vector<int32_t>& pos = g.pos[ofam(o)];
    std::vector<uint32_t>& out = per[(size_t)o];
    for (int k = -1; k <= g.Nz + 2; ++k)
      for (int j = -1; j <= g.Ny + 2; ++j)
        for (int i = -1; i <= g.Nx + 2; ++i) {
          const int64_t c = g.cell(i, j, k);
          const int64_t p = pos[(size_t)c];
          if (p == g.nstore) continue;
          const int oi = osx(o) > 0 ? i - 1 : g.Nx - i;
          const int oj = osy(o) > 0 ? j - 1 : g.Ny - j;
          const int ok = osz(o) > 0 ? k - 1 : g.Nz - k;
          const bool inside = i >= 1 && i <= g.Nx && j >= 1 && j <= g.Ny && k >= 1 && k <= g.Nz;
          if (inside && oi >= 1 && oj >= 1 && ok >= 1 && g.lab[(size_t)c] == FLUID) continue;
          out.push_back((uint32_t)p | (uint32_t)o << 28);
        }
    std::sort(out.begin(), out.end());
  });
  std::vector<uint32_t> all;
  for (auto& v : per) all.insert(all.end(), v.begin(), v.end());
  return all;
}

void build_wall_positions(Geometry& g, int threads) {
  const int64_t ns = g.nstore;
  // storage positions the wall kernels read and write
  const size_t nwall = g.wall_cell.size();
  const int K = Geometry::kWallPos;
  g.wall_pos.assign(nwall * K, (int32_t)ns);
  const int64_t stride[3] = {1, g.Nxt, g.Nxyt};
  parallel_for((int)((nwall + 4095) / 4096), threads, [&](int chunk) {
    for (size_t w = (size_t)chunk * 4096; w < std::min(nwall, (size_t)(chunk + 1) * 4096); ++w) {
      const int64_t c = g.wall_cell[w];
      int32_t* q = &g.wall_pos[w * K];
      for (int F = 0; F < 4; ++F) q[F] = g.pos[F][(size_t)c];
      const WallProg P = wall_prog(g.wall_label[w]);
      for (int fa = 0; fa < P.nface; ++fa)
        for (int d = 1; d <= 2; ++d)
          for (int k = 0; k < 4; ++k) {
            const int64_t n = c + (int64_t)d * P.face[fa].dir * stride[(int)P.face[fa].axis];
            q[4 + fa * 8 + (d - 1) * 4 + k] = g.pos[ofam(P.face[fa].in[k])][(size_t)n];
          }
    }
  });
}

int64_t flow_switched_word(const Geometry& g, int o, int64_t n, int axis) {
  if (g.lab[(size_t)n] < 20) return -1;
  const int w = g.wid[(size_t)n];
  const int src = (int)((g.wall_views[(size_t)w] >> (8 * o + (axis == 2 ? 4 : 0))) & 15u);
  if (src == 0) return -1;
  if (src <= 8) return (int64_t)(uint32_t)g.pos[ofam(src - 1)][(size_t)n] | (int64_t)((src - 1) ^ o) << 28;
  const int64_t corner = g.wall_corner[(size_t)w];
  if (corner < 0 || corner * 8 + o >= (1 << 28)) return -2;
  return (corner * 8 + o) | 0x90000000ll;
}

void make_window_relative(Geometry& g, int nc8, int threads) {
  parallel_for(g.nhp, threads, [&](int h) {
    for (int64_t t = g.hp_off[(size_t)h]; t < g.hp_off[(size_t)h + 1]; ++t) {
      int32_t* w = &g.hp_task[(size_t)(t * Geometry::kTaskWords)];
      const int o = (int)((uint32_t)w[0] >> 28);
      const int64_t lo = g.hp_lo[(size_t)h * 8 + o], len = g.hp_len[(size_t)h * 8 + o];
      auto rel = [&](int64_t p) { return (p == g.nstore ? len : p - lo) * nc8; };
      w[0] = (int32_t)((uint32_t)rel((uint32_t)w[0] & 0x0fffffffu) | (uint32_t)o << 28);
      for (int q = 1; q <= 6; ++q) w[q] = (int32_t)rel(w[q]);
    }
  });
}

int build_flow_records(const Geometry& g, int nc8, std::vector<int32_t>& rec, int64_t off[9], std::string& err,
                       int strips) {
  const int64_t ns = g.nstore;
  // storage index of a window-relative element offset (the zero slot: ns)
  auto slot_of = [&](int64_t h, int o, uint32_t rel) -> uint32_t {
    const int64_t lo = g.hp_lo[(size_t)h * 8 + o], len = g.hp_len[(size_t)h * 8 + o];
    return (int64_t)rel == len * nc8 ? (uint32_t)ns : (uint32_t)(rel / nc8 + lo);
  };
  std::vector<uint8_t> swept((size_t)8 * (ns + 1), 0);
  for (int o = 0; o < 8; ++o)
    for (int h = 0; h < g.nhp; ++h)
      for (int64_t t = g.hp_oct[(size_t)h * 8 + o]; t < g.hp_oct[(size_t)h * 8 + o + 1]; ++t) {
        const int32_t* w = &g.hp_task[(size_t)(t * Geometry::kTaskWords)];
        swept[(size_t)o * (ns + 1) + slot_of(h, o, (uint32_t)w[0] & 0x0fffffffu)] = 1;
      }
  rec.assign((size_t)g.hp_off[(size_t)g.nhp] * Geometry::kTaskWords, 0);
  if (strips < 1) strips = 1;
  const int64_t nyt = g.Nxyt / g.Nxt, band = (nyt + strips - 1) / strips;
  std::vector<std::pair<int, int64_t>> order;  // (hyperplane, task)
  std::vector<int64_t> first((size_t)strips + 1);
  int64_t at = 0;
  for (int o = 0; o < 8; ++o) {
    off[o] = at;
    const uint8_t* sw = &swept[(size_t)o * (ns + 1)];
    // strip of a task: its cell's band of y, counted along the octant's y
    // direction (upwind neighbours lie in the same band or an earlier one)
    auto strip_of = [&](int64_t t) {
      const int64_t cell = (uint32_t)g.hp_task[(size_t)(t * Geometry::kTaskWords) + 7] & 0x0fffffffu;
      const int64_t y = (cell % g.Nxyt) / g.Nxt;
      return (int)((osy(o) > 0 ? y : nyt - 1 - y) / band);
    };
    // stable counting sort of the hyperplane order by strip
    std::fill(first.begin(), first.end(), 0);
    for (int h = 0; h < g.nhp; ++h)
      for (int64_t t = g.hp_oct[(size_t)h * 8 + o]; t < g.hp_oct[(size_t)h * 8 + o + 1]; ++t)
        ++first[(size_t)strip_of(t) + 1];
    for (int s = 0; s < strips; ++s) first[(size_t)s + 1] += first[(size_t)s];
    order.assign((size_t)first[(size_t)strips], {0, 0});
    for (int h = 0; h < g.nhp; ++h)
      for (int64_t t = g.hp_oct[(size_t)h * 8 + o]; t < g.hp_oct[(size_t)h * 8 + o + 1]; ++t)
        order[(size_t)first[(size_t)strip_of(t)]++] = {h, t};
    for (const auto& ht : order) {
      const int h = ht.first;
      const int64_t t = ht.second;
      const int32_t* w = &g.hp_task[(size_t)(t * Geometry::kTaskWords)];
      int32_t* r = &rec[(size_t)(at++ * Geometry::kTaskWords)];
      r[0] = (int32_t)(slot_of(h, o, (uint32_t)w[0] & 0x0fffffffu) | (uint32_t)o << 28);
      uint32_t nbs[6];
      for (int q = 0; q < 6; ++q) nbs[q] = slot_of(h, o, (uint32_t)w[1 + q]);
      for (int ax = 0; ax < 3; ++ax) {
        const uint32_t d1 = nbs[2 * ax], d2 = nbs[2 * ax + 1];
        if (sw[d1]) nbs[2 * ax] |= 1u << 31;
        else if (sw[d2]) nbs[2 * ax + 1] |= 1u << 31;
      }
      uint32_t w7 = (uint32_t)w[7];
      if (w7 >> 31) {  // reads wall slots the reference switched: resolve them here
        const int64_t cell = w7 & 0x0fffffffu;
        const int64_t dy = (int64_t)osy(o) * g.Nxt, dz = (int64_t)osz(o) * g.Nxyt;
        const int64_t nbr[4] = {cell - dy, cell - 2 * dy, cell - dz, cell - 2 * dz};
        for (int q = 0; q < 4; ++q) {
          const int64_t nq = nbr[q];
          if (nq < 0 || nq >= g.Ntotal) {
            err = "single-launch sweep: a switched read outside the grid";
            return PIKS_E_GEOM;
          }
          const int64_t wq = flow_switched_word(g, o, nq, q < 2 ? 1 : 2);
          if (wq == -1) continue;
          if (wq < 0 || (nbs[2 + q] >> 31)) {
            err = wq < 0 ? "single-launch sweep: a switched read of a missing corner entry"
                         : "single-launch sweep: a switched read of a swept slot";
            return PIKS_E_GEOM;
          }
          nbs[2 + q] = (uint32_t)wq;
        }
        w7 &= 0x7fffffffu;
      }
      for (int q = 0; q < 6; ++q) r[1 + q] = (int32_t)nbs[q];
      r[7] = (int32_t)w7;
    }
  }
  off[8] = at;
  return 0;
}

}  // namespace piks
