// Host check of the single-launch sweep's records (build_flow_records) on one
// image, against the hyperplane schedule they are made from.
//
//   flow_records_check <raw image> <nx> <ny> <nz> <fluid layers> <wall order> [plant] [strips=N]
//
// With "plant" the records are given two mistakes before the checks (one
// task's wait mark cleared, one switched read pointed at the next slot),
// which the checks must report.  "strips=N" builds the records strip by
// strip (N bands of y).
// The schedule is built by build_geometry; a copy of its task records before
// make_window_relative (storage indices) is the reference.  Checks:
//  * every task of octant o appears once (in hyperplane order unless the
//    records go strip by strip);
//  * each axis waits on the nearest upwind neighbour the octant sweeps
//    (the first of the two that is swept), and only on it;
//  * every upwind source of the octant's own storage that the octant sweeps
//    is updated by an earlier task of the octant (so no wait is circular);
//  * every read of a switched wall slot addresses what the hyperplane
//    kernels' switched_source addresses (the views as stated there), every
//    other source the plain slot;
//  * the static split over W waves per octant drains (simulated round by
//    round: a wave finishes its current task once every source it waits on
//    is done), for W = 8, 64 and 768.
// Prints one line per check and "bad N"; exit status 1 if N > 0.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iterator>
#include <string>
#include <vector>

#include "geometry.h"
#include "wallprog.h"

using namespace piks;

int main(int argc, char** argv) {
  if (argc < 7 || argc > 9) {
    std::fprintf(stderr, "usage: %s raw nx ny nz layers wallorder [plant] [strips=N]\n", argv[0]);
    return 2;
  }
  bool plant = false;
  int strips = 1;
  for (int a = 7; a < argc; ++a) {
    const std::string s = argv[a];
    if (s == "plant") plant = true;
    else if (s.rfind("strips=", 0) == 0) strips = std::atoi(s.c_str() + 7);
    else {
      std::fprintf(stderr, "unknown argument %s\n", argv[a]);
      return 2;
    }
  }
  std::ifstream fh(argv[1], std::ios::binary);
  std::vector<uint8_t> raw((std::istreambuf_iterator<char>(fh)), {});
  GeomInput in{std::atoi(argv[2]), std::atoi(argv[3]), std::atoi(argv[4]), std::atoi(argv[5]), std::atoi(argv[6])};
  Geometry g;
  std::string err;
  if (build_geometry(in, raw.data(), (int64_t)raw.size(), g, err)) {
    std::printf("geometry: %s\n", err.c_str());
    return 2;
  }
  const int nc8 = 64;
  const int64_t ns = g.nstore, ntask = g.hp_off[(size_t)g.nhp];
  const std::vector<int32_t> plain = g.hp_task;  // storage indices
  make_window_relative(g, nc8, 4);
  std::vector<int32_t> rec;
  int64_t off[9];
  if (build_flow_records(g, nc8, rec, off, err, strips)) {
    std::printf("records: %s\n", err.c_str());
    return 1;
  }
  if (plant) {
    bool wait_done = false, switch_done = false;
    for (int64_t t = 0; t < off[8] && !(wait_done && switch_done); ++t) {
      int32_t* r = &rec[(size_t)t * 8];
      for (int q = 0; q < 6; ++q) {
        const uint32_t wq = (uint32_t)r[1 + q];
        if (!wait_done && (wq >> 28) == 8u) {
          r[1 + q] = (int32_t)(wq & 0x0fffffffu);
          wait_done = true;
        } else if (!switch_done && q >= 2 && ((wq >> 28) & 7u) != 0u) {
          r[1 + q] = (int32_t)(wq + 1u);
          switch_done = true;
        }
      }
    }
    std::printf("planted: wait mark %d, switched read %d\n", (int)wait_done, (int)switch_done);
  }
  int64_t bad = 0, waits = 0, switched = 0, other_octant = 0, corner = 0;
  auto fail = [&](const char* what, int o, int64_t k) {
    if (bad++ < 10) std::printf("FAIL %s (octant %d, task %lld)\n", what, o, (long long)k);
  };
  if (off[8] != ntask) fail("task count", -1, off[8]);
  std::vector<std::vector<int64_t>> prod(8, std::vector<int64_t>((size_t)ns + 1, -1));
  std::vector<std::vector<std::vector<int64_t>>> deps(8);
  for (int o = 0; o < 8; ++o) {
    // the octant's tasks in the records' order: each record's plain task (the
    // one that writes the same slot), every plain task of the octant once
    std::vector<int64_t> order, plain_order, task_of((size_t)ns + 1, -1);
    std::vector<uint8_t> swept((size_t)ns + 1, 0), seen((size_t)ns + 1, 0);
    for (int h = 0; h < g.nhp; ++h)
      for (int64_t t = g.hp_oct[(size_t)h * 8 + o]; t < g.hp_oct[(size_t)h * 8 + o + 1]; ++t) {
        const uint32_t s = (uint32_t)plain[(size_t)t * 8] & 0x0fffffffu;
        plain_order.push_back(t);
        swept[(size_t)s] = 1;
        task_of[(size_t)s] = t;
      }
    if ((int64_t)plain_order.size() != off[o + 1] - off[o]) fail("octant task count", o, (int64_t)plain_order.size());
    for (int64_t k = 0; k < off[o + 1] - off[o]; ++k) {
      const uint32_t s = (uint32_t)rec[(size_t)(off[o] + k) * 8] & 0x0fffffffu;
      if (s >= (uint32_t)ns || task_of[s] < 0 || seen[s]) {
        fail("slot not a task of the octant, or twice", o, k);
        continue;
      }
      seen[s] = 1;
      order.push_back(task_of[s]);
    }
    if (order.size() != plain_order.size()) {
      deps[o].resize((size_t)(off[o + 1] - off[o]));
      continue;
    }
    if (strips == 1 && order != plain_order) fail("hyperplane order", o, 0);
    for (size_t k = 0; k < order.size(); ++k)
      prod[o][(size_t)((uint32_t)plain[order[k] * 8] & 0x0fffffffu)] = (int64_t)k;
    deps[o].resize(order.size());
    for (size_t k = 0; k < order.size(); ++k) {
      const int32_t* p = &plain[(size_t)order[k] * 8];
      const int32_t* r = &rec[(size_t)(off[o] + (int64_t)k) * 8];
      if (((uint32_t)r[0] & 0x0fffffffu) != ((uint32_t)p[0] & 0x0fffffffu)) fail("slot", o, (int64_t)k);
      const uint32_t w7 = (uint32_t)p[7], cell = w7 & 0x0fffffffu;
      if (((uint32_t)r[7] & 0x7fffffffu) != (w7 & 0x7fffffffu) || ((uint32_t)r[7] >> 31)) fail("word 7", o, (int64_t)k);
      // the sources the hyperplane kernels read: plain slots, switched where
      // the task's record flags it (view_fix: y neighbours at distance 1, 2,
      // then z)
      const int64_t dy = (int64_t)osy(o) * g.Nxt, dz = (int64_t)osz(o) * g.Nxyt;
      const int64_t nbr[4] = {(int64_t)cell - dy, (int64_t)cell - 2 * dy, (int64_t)cell - dz,
                              (int64_t)cell - 2 * dz};
      for (int q = 0; q < 6; ++q) {
        const uint32_t want_slot = (uint32_t)p[1 + q];
        const uint32_t wq = (uint32_t)r[1 + q];
        const uint32_t x = (wq >> 28) & 7u, idx = wq & 0x0fffffffu;
        // the switched source, stated as switched_source states it
        int kind = 0;  // 0 plain, 1 another octant's f, 2 corner store
        int64_t src_oct = o, src_idx = want_slot;
        if (q >= 2 && (w7 >> 31)) {
          const int64_t n = nbr[q - 2];
          if (g.lab[(size_t)n] >= 20) {
            const int wi = g.wid[(size_t)n];
            const int src = (int)((g.wall_views[(size_t)wi] >> (8 * o + (q < 4 ? 0 : 4))) & 15u);
            if (src != 0 && src <= 8) {
              kind = 1;
              src_oct = src - 1;
              src_idx = g.pos[ofam(src - 1)][(size_t)n];
            } else if (src > 8) {
              kind = 2;
              src_idx = (int64_t)g.wall_corner[(size_t)wi] * 8 + o;
            }
          }
        }
        if (kind == 0) {
          if (x != 0 || idx != want_slot) fail("plain source", o, (int64_t)k);
        } else if (kind == 1) {
          ++switched;
          other_octant += src_oct != o;
          if ((wq >> 31) || (int64_t)(o ^ x) != src_oct || idx != src_idx) fail("switched f source", o, (int64_t)k);
        } else {
          ++switched;
          ++corner;
          if (!(wq >> 31) || x == 0 || (int64_t)idx != src_idx || src_idx >= (int64_t)g.ncorner * 8)
            fail("switched corner source", o, (int64_t)k);
        }
        // own-octant swept sources are updated earlier
        if (kind == 0 && swept[want_slot]) {
          const int64_t pk = prod[o][want_slot];
          if (pk < 0 || pk >= (int64_t)k) fail("source updated later", o, (int64_t)k);
        }
      }
      // the waits: per axis the first of the two neighbours the octant sweeps
      for (int ax = 0; ax < 3; ++ax) {
        const uint32_t d1 = (uint32_t)p[1 + 2 * ax], d2 = (uint32_t)p[2 + 2 * ax];
        const int want = swept[d1] ? 2 * ax : swept[d2] ? 2 * ax + 1 : -1;
        int got = -1, marks = 0;
        for (int q = 2 * ax; q < 2 * ax + 2; ++q)
          if (((uint32_t)r[1 + q] >> 28) == 8u) {
            got = q;
            ++marks;
          }
        if (marks > 1 || got != want) fail("wait mark", o, (int64_t)k);
        if (want >= 0) {
          ++waits;
          deps[o][k].push_back(prod[o][want == 2 * ax ? d1 : d2]);
        }
      }
    }
  }
  std::printf("tasks %lld, waits %lld, switched reads %lld (another octant %lld, corner store %lld)\n",
              (long long)ntask, (long long)waits, (long long)switched, (long long)other_octant, (long long)corner);
  // the static split drains
  for (int W : {8, 64, 768}) {
    int64_t rounds_max = 0;
    for (int o = 0; o < 8; ++o) {
      const int64_t n = off[o + 1] - off[o];
      std::vector<uint8_t> done((size_t)n, 0);
      std::vector<int64_t> cur((size_t)W);
      for (int w = 0; w < W; ++w) cur[(size_t)w] = w;
      int64_t left = n, rounds = 0;
      while (left > 0) {
        ++rounds;
        std::vector<int64_t> fin;
        for (int w = 0; w < W; ++w) {
          const int64_t k = cur[(size_t)w];
          if (k >= n) continue;
          bool ready = true;
          for (int64_t d : deps[o][(size_t)k]) ready = ready && done[(size_t)d];
          if (ready) fin.push_back(w);
        }
        if (fin.empty()) {
          fail("static split stalls", o, left);
          break;
        }
        for (int64_t w : fin) {
          done[(size_t)cur[(size_t)w]] = 1;
          cur[(size_t)w] += W;
          --left;
        }
      }
      rounds_max = std::max(rounds_max, rounds);
    }
    std::printf("W %d: drains in at most %lld rounds\n", W, (long long)rounds_max);
  }
  std::printf("bad %lld\n", (long long)bad);
  return bad != 0;
}
