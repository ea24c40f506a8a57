// Goldstein branch-cut placement for phase unwrapping, with the reference's
// ordering and search rules (Codes/sub_code/goldstein_GPU/goldstein_branch_cut.m:1-87).
//
// Residues of a column-major (m1, m2, n3) map are listed as rows of six int32
// [flat index, charge, case, d1, d2, z], sorted by (z, case) keeping flat-index
// order inside a case.  Each z is split into sq x sq search cases of side
// ceil(m / sq), sq = ceil(sqrt((count + 1) / 16)); the lookup table holds, per z
// and case (d1 case fastest), the first listed residue of the case or -1.
//
// The nearest-residue box search is shared by the device pass (8 nearest of every
// residue, no state test) and the host branch search (1 nearest residue that is not
// active).  Distances are int(sqrtf((float)(dx*dx + dy*dy))); the float results are
// formed from double arithmetic rounded once to float, which equals the correctly
// rounded single-precision operation for a quotient or square root, so host and
// device give the same integers.
#pragma once
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

namespace cbs {
namespace gs {

constexpr int kCols = 6;
constexpr int kIdx = 0, kVal = 1, kCase = 2, kX = 3, kY = 4, kZ = 5;
constexpr int kNearest = 8;            // goldstein_branch_cut.m:19
constexpr unsigned char kBalanced = 1, kActive = 2;

__host__ __device__ inline int imax(int a, int b) { return a > b ? a : b; }
__host__ __device__ inline int imin(int a, int b) { return a < b ? a : b; }
__host__ __device__ inline int iabs(int a) { return a < 0 ? -a : a; }

__host__ __device__ inline int dist_of(int dx, int dy) {
  const float n = (float)(dx * dx + dy * dy);
  return (int)(float)sqrt((double)n);
}

__host__ __device__ inline int case_side(int m, int sq) { return (int)ceil((double)(float)((double)m / (double)sq)); }

// search_universal.cpp:24-49: insertion point of x among the P sorted distances.
__host__ __device__ inline int insert_pos(const int* arr, int x, int low, int high) {
  int mid;
  while (low < high) {
    mid = (high + low) / 2;
    if (arr[2 * mid + 1] == x) break;
    if (arr[2 * mid + 1] > x)
      high = mid - 1;
    else
      low = mid + 1;
  }
  mid = (high + low) / 2;
  return x <= arr[2 * mid + 1] ? mid : mid + 1;
}

// search_universal.cpp:51-114: fold the residues of case (b1, b2) into the list.
template <int P>
__host__ __device__ inline void search_case(const unsigned char* state, const int32_t* table, int table_size,
                                            const int32_t* rows, int n, int* cur, int sq, int b1, int b2, int zst,
                                            int x0, int y0, int self) {
  if (b1 >= sq || b2 >= sq) return;
  const int look = zst + b1 + b2 * sq;
  // the ring bound keeps every searched box inside the table (a box before the first
  // case is never closer than the border); a negative id is skipped all the same
  if (look < 0 || look >= table_size) return;
  const int start = table[look];
  if (start == -1) return;
  int end = n;
  for (int k = 1; look + k < table_size; ++k) {
    const int v = table[look + k];
    if (v >= 0) {
      end = v;
      break;
    }
  }
  for (int i = start; i < end; ++i) {
    if (state && (state[i] & kActive)) continue;
    const int dx = iabs(x0 - rows[kCols * i + kX]);
    const int dy = iabs(y0 - rows[kCols * i + kY]);
    const int d = dist_of(dx, dy);
    if (d < cur[2 * P - 1] && i != self) {
      const int ins = insert_pos(cur, d, 0, P - 1);
      if (ins < P) {
        for (int k = P - 1; k > ins; --k) {
          cur[2 * k] = cur[2 * (k - 1)];
          cur[2 * k + 1] = cur[2 * (k - 1) + 1];
        }
        cur[2 * ins] = i;
        cur[2 * ins + 1] = d;
      }
    }
  }
}

// search_universal.cpp:116-271: the P nearest residues of listed residue `self` in its z,
// as [id, distance] pairs; unfilled entries are (-1, distance to the border).
template <int P>
__host__ __device__ inline void nearest(const unsigned char* state, const int32_t* rows, int n, const int32_t* table,
                                        int table_size, const int32_t* zstart, const int32_t* square, int s1, int s2,
                                        int self, int* cur) {
  const int* r = rows + kCols * self;
  const int cs_id = r[kCase], x0 = r[kX], y0 = r[kY], z0 = r[kZ];
  const int zst = zstart[z0];
  const int sq = square[z0];
  const int cs1 = case_side(s1, sq), cs2 = case_side(s2, sq);
  const int c1 = cs_id % sq, c2 = cs_id / sq;
  const int max_border = imax(imax(sq - c1, sq - c2), imax(c1, c2));
  const int dmax = imin(imin(x0, y0), imin(s1 - x0, s2 - y0));
  for (int k = 0; k < P; ++k) {
    cur[2 * k] = -1;
    cur[2 * k + 1] = dmax;
  }
  int last = dmax;
  for (int d = 0; d < max_border && d < sq; ++d) {
    if (d == 0) {
      search_case<P>(state, table, table_size, rows, n, cur, sq, c1, c2, zst, x0, y0, self);
    } else {
      for (int ii = 1 - d; ii <= d; ++ii) {
        const int nz = ii == 0 ? 0 : 1;
        const int bx[4] = {c1 - d, c1 + ii, c1 - ii, c1 + d};
        const int by[4] = {c2 - ii, c2 - d, c2 + d, c2 + ii};
        const int md[4] = {
            imax(iabs((bx[0] + 1) * cs1 - x0), iabs(((by[0] + (ii > 0 ? 1 : 0)) * cs2 - y0) * nz)),
            imax(iabs(((bx[1] + (ii > 0 ? 0 : 1)) * cs1 - x0) * nz), iabs((by[1] + 1) * cs2 - y0)),
            imax(iabs(((bx[2] + (ii > 0 ? 1 : 0)) * cs1 - x0) * nz), iabs(by[2] * cs2 - y0)),
            imax(iabs(bx[3] * cs1 - x0), iabs(((by[3] + (ii > 0 ? 0 : 1)) * cs2 - y0) * nz))};
        for (int e = 0; e < 4; ++e)
          if (md[e] < last) search_case<P>(state, table, table_size, rows, n, cur, sq, bx[e], by[e], zst, x0, y0, self);
        last = cur[2 * P - 1];
      }
    }
    last = cur[2 * P - 1];
    if (d * cs1 >= last && d * cs2 >= last) break;
  }
}

// ---- host part: listing and the sequential branch search

// goldstein_branch_cut.m:7-38: list, order and index the residues of a dipole-free map.
struct GsList {
  int n = 0;
  std::vector<int32_t> rows;                  // 6 per residue, see goldstein.hpp
  std::vector<int32_t> square, zstart, table;
};

inline GsList gs_list(const std::vector<float>& res, int m1, int m2, int n3) {
  GsList L;
  const long long plane = (long long)m1 * m2;
  std::vector<long long> idx;
  std::vector<int> count(n3, 0);
  for (long long i = 0; i < (long long)res.size(); ++i)
    if (res[i] != 0.0f) {
      idx.push_back(i);
      ++count[(int)(i / plane)];
    }
  L.square.resize(n3);
  L.zstart.resize(n3);
  int size = 0;
  for (int z = 0; z < n3; ++z) {
    L.square[z] = (int32_t)std::ceil(std::sqrt((count[z] + 1) / (double)(2 * kNearest)));
    L.zstart[z] = size;
    size += L.square[z] * L.square[z];
  }
  L.n = (int)idx.size();
  std::vector<int32_t> rows(6 * (size_t)L.n);
  for (int k = 0; k < L.n; ++k) {
    const long long i = idx[k];
    const int z = (int)(i / plane), d12 = (int)(i % plane);
    const int d1 = d12 % m1, d2 = d12 / m1;
    const int sq = L.square[z];
    const int c = d1 / case_side(m1, sq) + (d2 / case_side(m2, sq)) * sq;
    int32_t* r = &rows[6 * (size_t)k];
    r[kIdx] = (int32_t)i;
    r[kVal] = (int32_t)std::lrint(res[i]);
    r[kCase] = c;
    r[kX] = d1;
    r[kY] = d2;
    r[kZ] = z;
  }
  std::vector<int> order(L.n);
  for (int k = 0; k < L.n; ++k) order[k] = k;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    const int za = rows[6 * (size_t)a + kZ], zb = rows[6 * (size_t)b + kZ];
    if (za != zb) return za < zb;
    return rows[6 * (size_t)a + kCase] < rows[6 * (size_t)b + kCase];
  });
  L.rows.resize(rows.size());
  for (int k = 0; k < L.n; ++k) std::copy_n(&rows[6 * (size_t)order[k]], 6, &L.rows[6 * (size_t)k]);
  L.table.assign(std::max(size, 1), -1);
  for (int k = 0; k < L.n; ++k) {
    const int32_t* r = &L.rows[6 * (size_t)k];
    const int32_t* prev = r - 6;
    if (k == 0 || r[kCase] != prev[kCase] || r[kZ] != prev[kZ]) L.table[L.zstart[r[kZ]] + r[kCase]] = k;
  }
  return L;
}

// get_branches_goldstein_compute_v2.cpp:16-203 and get_branches_goldstein_update_v2.cpp:12-60:
// from every residue not yet balanced, grow a tree over the residues, extending at each
// step an active node whose next candidate is at the current smallest distance, until
// the tree's charge is zero or a node is joined to the border.  sl: the device nearest
// lists, [(i * 8 + k) * 2 + {0, 1}].  Returns 6 int32 per branch
// [p1 d1, p1 d2, p2 d1, p2 d2, z, jump], p1 = (-1, -1) for the border.
inline std::vector<int32_t> gs_branches(const GsList& L, const std::vector<int32_t>& sl, int m1, int m2) {
  constexpr int P = kNearest;
  constexpr int kDistMax = 200000;
  struct Node {
    int residue, parent, parent_branch, min_distance, min_idx, overflow_id;
  };
  const int n = L.n;
  std::vector<unsigned char> state(n, 0);
  std::vector<int32_t> out;
  std::vector<Node> act;
  auto row = [&](int i, int c) { return L.rows[6 * (size_t)i + c]; };
  auto shortest_id = [&](const Node& nd) {
    return nd.min_idx < P ? sl[(size_t)nd.residue * 2 * P + nd.min_idx * 2] : nd.overflow_id;
  };
  auto update = [&](Node& nd) {
    for (int nx = 1 + nd.min_idx;; ++nx) {
      if (nx < P) {
        const int cand = sl[(size_t)nd.residue * 2 * P + nx * 2];
        if (cand == -1 || nx == P - 1 || !(state[cand] & kActive)) {
          nd.min_idx = nx;
          nd.min_distance = sl[(size_t)nd.residue * 2 * P + nx * 2 + 1];
          return;
        }
      } else {
        nd.min_idx = P;
        int cur[2];
        nearest<1>(state.data(), L.rows.data(), n, L.table.data(), (int)L.table.size(), L.zstart.data(),
                       L.square.data(), m1, m2, nd.residue, cur);
        nd.overflow_id = cur[0];
        nd.min_distance = cur[1];
        return;
      }
    }
  };
  for (int i = 0; i < n; ++i) {
    if (state[i] & kBalanced) continue;
    act.clear();
    act.push_back(Node{i, 0, -1, 0, 0, -1});
    state[i] |= kActive;
    int total = row(i, kVal);
    int next_d = 0;
    while (total != 0) {
      int cur_d = next_d;
      next_d = kDistMax;
      for (size_t a = 0; a < act.size(); ++a) {
        if (act[a].min_distance == cur_d) {
          const int sid = shortest_id(act[a]);
          if (sid == -1 || !(state[sid] & kActive)) {
            const int old = total;
            if (sid != -1 && !(state[sid] & kBalanced)) total += row(sid, kVal);
            if (sid == -1) total = 0;
            act.push_back(Node{sid, (int)a, (int)(out.size() / 6), 0, -1, -1});
            if (sid != -1) state[sid] |= kActive;
            const int r = act[a].residue;
            const int32_t br[6] = {sid != -1 ? row(sid, kX) : -1, sid != -1 ? row(sid, kY) : -1,
                                   row(r, kX), row(r, kY), row(r, kZ), 0};
            out.insert(out.end(), br, br + 6);
            if (sid == -1 || !(state[sid] & kBalanced)) {
              const int jump = sid != -1 ? row(sid, kVal) : -old;
              for (int p = (int)act.size() - 1; p != 0; p = act[p].parent) out[6 * (size_t)act[p].parent_branch + 5] += jump;
            }
            update(act[a]);
            if (sid != -1) update(act.back());
            // the reference compares the distance of the node as pushed, which is 0
            cur_d = std::min(cur_d, 0);
            next_d = std::min(next_d, 0);
          } else {
            update(act[a]);
          }
        }
        if (act[a].min_distance < next_d) next_d = act[a].min_distance;
        if (total == 0) break;
      }
    }
    for (const Node& nd : act)
      if (nd.residue != -1) state[nd.residue] = (unsigned char)((state[nd.residue] & ~kActive) | kBalanced);
  }
  return out;
}

}  // namespace gs
}  // namespace cbs
