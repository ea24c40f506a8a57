"""Goldstein branch-cut phase unwrapping as the reference runs it (restatement).

Test infrastructure only (see oracle/__init__.py).

Follows Codes/sub_code/goldstein_GPU/unwrapp2_gpu.m:1-5 and goldstein_branch_cut.m:1-87:
- residues (oracle.rytov.residues, get_residue_cuda.cu:35-66);
- dipole removal, four directions in turn (remove_dipoles_cuda.cu:50-106,122-130);
- residues ordered by (z, search case), stable (goldstein_branch_cut.m:16-27,
  create_sort_value_cuda.cu:60-86);
- the case lookup table (get_lookup_mex.cpp:128-135, get_lookup_cuda.cu:63-89);
- the 8 nearest residues of every residue by the box search
  (search_universal.cpp:19-271 compiled with GPU_COMPUTE, get_shortest_sort_cuda.cu:64-87);
- the branch search, sequential (get_branches_goldstein_compute_v2.cpp:16-203,
  get_branches_goldstein_update_v2.cpp:24-60, search_universal.cpp without GPU_COMPUTE);
- rastering the branches into the 2 pi step maps (raster_branches_cuda.cu:65-135);
- the raster integration (oracle.rytov.raster_unwrap).

Coordinates: residue maps are (m1, m2, n3) = (n1 - 1, n2 - 1, n3), column-major, and the
flat index is d1 + m1 * d2 + m1 * m2 * z, as MATLAB's find() numbers them.  C integer
division truncates toward zero and is written ``_cdiv`` where an operand can be negative.
"""
from __future__ import annotations

import math

import numpy as np

from .rytov import residues, raster_unwrap

PRECOMPUTED = 8          # goldstein_branch_cut.m:19 (precomputed_shortest_number)
DIST_MAX = 200000        # search_universal.cpp:147, get_branches_goldstein_compute_v2.cpp:18
ACTIVE, BALANCED = 2, 1  # get_branches_goldstein_update_v2.hpp:10-11
COL_IDX, COL_VAL, COL_CASE, COL_X, COL_Y, COL_Z = range(6)


def _cdiv(a, b):
    q = abs(a) // abs(b)
    return q if (a >= 0) == (b > 0) else -q


def remove_dipoles(res):
    """remove_dipoles_cuda.cu:50-106 over the positive residues (find order), directions
    0..3 = d1 - 1, d2 - 1, d1 + 1, d2 + 1, one launch each.  A positive residue next to a
    negative one in the current direction cancels min(+, -) of both; a dipole along d2
    writes +-removed into step1 at (d1 + 1, min(d2, d2') + 1) (directions 1 and 3 only).
    Returns (residues, step1 (n1, n2, n3) float32, step2 (n2, n3) float32 zeros)."""
    m1, m2, n3 = np.shape(res)
    n1, n2 = m1 + 1, m2 + 1
    flat = np.asarray(res, np.float32).reshape(-1, order='F').copy()
    s1 = np.zeros(n1 * n2 * n3, np.float32)
    pos = (np.flatnonzero(flat > 0) + 1).astype(np.int64)     # int32(find(residue_map(:) > 0))
    for direction in range(4):
        # one launch: no two threads touch the same element (each negative neighbour
        # belongs to one positive per direction), so program order gives the same result
        for k in range(len(pos)):
            if pos[k] <= 0:
                continue
            rid = int(pos[k]) - 1
            d12, z = rid % (m1 * m2), rid // (m1 * m2)
            d2, d1 = d12 // m1, d12 % m1
            p1 = d1 - (1 if direction == 0 else 0) + (1 if direction == 2 else 0)
            p2 = d2 - (1 if direction == 1 else 0) + (1 if direction == 3 else 0)
            if not (0 <= p1 < m1 and 0 <= p2 < m2):
                continue
            pid = z * m1 * m2 + p2 * m1 + p1
            pv = flat[pid]
            if pv < 0:
                v = flat[rid]
                rem = np.float32(min(v, -pv))
                v = np.float32(np.round(np.float32(v - rem)))
                pv = np.float32(np.round(np.float32(pv + rem)))
                flat[rid] = v
                flat[pid] = pv
                if v == 0:
                    pos[k] = -1
                if direction in (1, 3):
                    c1, c2 = p1 + 1, min(d2, p2) + 1
                    s1[z * n1 * n2 + c2 * n1 + c1] = rem if direction == 1 else -rem
    return (flat.reshape((m1, m2, n3), order='F'), s1.reshape((n1, n2, n3), order='F'),
            np.zeros((n2, n3), np.float32, order='F'))


def order_residues(res):
    """goldstein_branch_cut.m:7-27: (6, N) int32 rows [flat index, value, case, d1, d2, z]
    sorted stably by z then case, and the per-z number of cases per side."""
    m1, m2, n3 = res.shape
    flat = res.reshape(-1, order='F')
    idx = np.flatnonzero(flat)
    z = idx // (m1 * m2)
    count = np.bincount(z, minlength=n3)[:n3]
    square = np.array([int(math.ceil(math.sqrt((c + 1) / (PRECOMPUTED * 2)))) for c in count], np.int32)
    d12 = idx % (m1 * m2)
    d1, d2 = d12 % m1, d12 // m1
    sq = square[z]
    cs1 = np.ceil(np.float32(m1) / sq.astype(np.float32)).astype(np.int64)   # create_sort_value_cuda.cu:79-80
    cs2 = np.ceil(np.float32(m2) / sq.astype(np.float32)).astype(np.int64)
    case = d1 // cs1 + (d2 // cs2) * sq
    val = np.round(flat[idx]).astype(np.int32)
    rows = np.stack([idx, val, case, d1, d2, z]).astype(np.int32)
    order = np.lexsort((case, z))                   # stable: ties keep find() order
    return np.ascontiguousarray(rows[:, order]), square


def lookup(ordered, square):
    """get_lookup_mex.cpp:128-135 and get_lookup_cuda.cu:63-89: per z a square x square
    table (case d1 fastest) of the first ordered residue in each case, -1 when empty."""
    zstart = np.zeros(len(square), np.int32)
    size = 0
    for i, s in enumerate(square):
        zstart[i] = size
        size += int(s) * int(s)
    table = np.full(size, -1, np.int32)
    n = ordered.shape[1]
    for i in range(n):
        c, z = ordered[COL_CASE, i], ordered[COL_Z, i]
        if i == 0 or c != ordered[COL_CASE, i - 1] or z != ordered[COL_Z, i - 1]:
            table[zstart[z] + c] = i
    return table, zstart


def _bsearch(arr, x, low, high):
    """search_universal.cpp:24-49 (binarySearch_array_2_size), C arithmetic."""
    while low < high:
        mid = _cdiv(high + low, 2)
        if arr[2 * mid + 1] == x:
            break
        if arr[2 * mid + 1] > x:
            high = mid - 1
        else:
            low = mid + 1
    mid = _cdiv(high + low, 2)
    return mid if x <= arr[2 * mid + 1] else mid + 1


def _search_case(states, table, ordered, cur, sq, b1, b2, zst, n, x0, y0, P, self_id):
    """search_universal.cpp:51-114 (search_in_case); states None = the GPU form (no
    state test)."""
    if b1 >= sq or b2 >= sq:
        return
    look = zst + b1 + b2 * sq
    if look < 0:
        raise IndexError('box search read before the lookup table')   # never reached (see module tests)
    start = table[look]
    if start == -1:
        return
    end = n
    k = 1
    while look + k < len(table) and end == n:
        v = table[look + k]
        if v >= 0:
            end = int(v)
            break
        k += 1
    for i in range(int(start), end):
        if states is not None and states[i] & ACTIVE:
            continue
        dx = abs(x0 - int(ordered[COL_X, i]))
        dy = abs(y0 - int(ordered[COL_Y, i]))
        dist = int(np.sqrt(np.float32(dx * dx + dy * dy)))       # int(sqrtf(float))
        if dist < cur[2 * P - 1] and i != self_id:
            ins = _bsearch(cur, dist, 0, P - 1)
            if ins < P:
                for kk in range(P - 1, ins, -1):
                    cur[2 * kk] = cur[2 * (kk - 1)]
                    cur[2 * kk + 1] = cur[2 * (kk - 1) + 1]
                cur[2 * ins] = i
                cur[2 * ins + 1] = dist


def nearest(ordered, table, zstart, square, size3, i, P=PRECOMPUTED, states=None):
    """search_universal.cpp:116-271 (shortest_compute_kernel) for ordered residue i:
    [id, distance] of the P nearest residues of its z, then (-1, distance to the border)."""
    n = ordered.shape[1]
    s1, s2 = int(size3[0]), int(size3[1])
    case = int(ordered[COL_CASE, i])
    x0, y0, z0 = int(ordered[COL_X, i]), int(ordered[COL_Y, i]), int(ordered[COL_Z, i])
    zst = int(zstart[z0])
    sq = int(square[z0])
    cs1 = int(math.ceil(np.float32(s1) / np.float32(sq)))
    cs2 = int(math.ceil(np.float32(s2) / np.float32(sq)))
    c1, c2 = case % sq, case // sq
    max_border = max(max(sq - c1, sq - c2), max(c1, c2))
    dmax = min(min(x0, y0), min(s1 - x0, s2 - y0))
    cur = [0] * (2 * P)
    for k in range(P):
        cur[2 * k], cur[2 * k + 1] = -1, dmax
    last = dmax
    d = 0
    while d < max_border and d < sq:
        if d == 0:
            _search_case(states, table, ordered, cur, sq, c1, c2, zst, n, x0, y0, P, i)
        else:
            for ii in range(1 - d, d + 1):
                nz = 0 if ii == 0 else 1
                bx = (c1 - d, c1 + ii, c1 - ii, c1 + d)
                by = (c2 - ii, c2 - d, c2 + d, c2 + ii)
                md = (max(abs((bx[0] + 1) * cs1 - x0), abs(((by[0] + (1 if ii > 0 else 0)) * cs2 - y0) * nz)),
                      max(abs(((bx[1] + (0 if ii > 0 else 1)) * cs1 - x0) * nz), abs((by[1] + 1) * cs2 - y0)),
                      max(abs(((bx[2] + (1 if ii > 0 else 0)) * cs1 - x0) * nz), abs(by[2] * cs2 - y0)),
                      max(abs(bx[3] * cs1 - x0), abs(((by[3] + (0 if ii > 0 else 1)) * cs2 - y0) * nz)))
                for e in range(4):
                    if md[e] < last:
                        _search_case(states, table, ordered, cur, sq, bx[e], by[e], zst, n, x0, y0, P, i)
                last = cur[2 * P - 1]
        last = cur[2 * P - 1]
        if d * cs1 >= last and d * cs2 >= last:
            break
        d += 1
    return cur


def shortest_list(ordered, table, zstart, square, size3, P=PRECOMPUTED):
    """get_shortest_sort_gpu: (2, P, N) int32, [id, distance] per rank per residue."""
    n = ordered.shape[1]
    out = np.empty((2, P, n), np.int32, order='F')
    for i in range(n):
        cur = nearest(ordered, table, zstart, square, size3, i, P)
        out[:, :, i] = np.array(cur, np.int32).reshape(P, 2).T
    return out


class _Node:
    __slots__ = ('residue', 'parent', 'parent_branch', 'min_distance', 'min_distance_idx', 'overflow_id')

    def __init__(self, residue, parent, min_distance, idx, parent_branch=-1):
        self.residue = residue
        self.parent = parent
        self.parent_branch = parent_branch
        self.min_distance = min_distance
        self.min_distance_idx = idx
        self.overflow_id = -1


def branches(ordered, short, square, table, zstart, size3):
    """get_branches_goldstein_compute_v2.cpp:16-203: grow a tree from every unbalanced
    residue in order, always extending the active node whose next nearest candidate is
    closest, until the charge balances or a node links to the border.  Returns (6, B)
    int32 rows [p1 d1, p1 d2, p2 d1, p2 d2, z, jump]; p1 = (-1, -1) is the border."""
    n = ordered.shape[1]
    P = short.shape[1]
    sl = short.reshape(-1, order='F')               # [residue * 2P + k * 2 + 0/1]
    state = np.zeros(n, np.uint8)
    out = []

    def shortest_id(nd):                            # get_branches_goldstein_update_v2.cpp:12-17
        if nd.min_distance_idx < P:
            return int(sl[nd.residue * 2 * P + nd.min_distance_idx * 2])
        return nd.overflow_id

    def update(nd):                                 # get_branches_goldstein_update_v2.cpp:24-60
        nxt = 1 + nd.min_distance_idx
        while True:
            if nxt < P:
                cand = int(sl[nd.residue * 2 * P + nxt * 2])
                if cand == -1 or nxt == P - 1 or not (state[cand] & ACTIVE):
                    nd.min_distance_idx = nxt
                    nd.min_distance = int(sl[nd.residue * 2 * P + nxt * 2 + 1])
                    return
            else:
                nd.min_distance_idx = P
                r = nearest(ordered, table, zstart, square, size3, nd.residue, P=1, states=state)
                nd.overflow_id, nd.min_distance = r[0], r[1]
                return
            nxt += 1

    for i in range(n):
        if state[i] & BALANCED:
            continue
        act = [_Node(i, 0, 0, 0)]
        state[i] |= ACTIVE
        total = int(ordered[COL_VAL, i])
        nxt_d = 0
        while total != 0:
            cur_d = nxt_d
            nxt_d = DIST_MAX
            a = 0
            while a < len(act):
                node = act[a]
                if node.min_distance == cur_d:
                    sid = shortest_id(node)
                    if sid == -1 or not (state[sid] & ACTIVE):
                        old = total
                        if sid != -1 and not (state[sid] & BALANCED):
                            total += int(ordered[COL_VAL, sid])
                        if sid == -1:
                            total = 0
                        new = _Node(sid, a, 0, -1, parent_branch=len(out))
                        act.append(new)
                        if sid != -1:
                            state[sid] |= ACTIVE
                        r = node.residue
                        br = [int(ordered[COL_X, sid]) if sid != -1 else -1,
                              int(ordered[COL_Y, sid]) if sid != -1 else -1,
                              int(ordered[COL_X, r]), int(ordered[COL_Y, r]), int(ordered[COL_Z, r]), 0]
                        out.append(br)
                        if sid == -1 or not (state[sid] & BALANCED):
                            jump = int(ordered[COL_VAL, sid]) if sid != -1 else -old
                            p = len(act) - 1
                            while p != 0:
                                out[act[p].parent_branch][5] += jump
                                p = act[p].parent
                        update(node)
                        if sid != -1:
                            update(new)
                        # the reference compares the pushed copy's min_distance, which is 0
                        cur_d = min(cur_d, 0)
                        nxt_d = min(nxt_d, 0)
                    else:
                        update(node)
                if node.min_distance < nxt_d:
                    nxt_d = node.min_distance
                if total == 0:
                    break
                a += 1
        for nd in act:
            if nd.residue != -1:
                state[nd.residue] = (int(state[nd.residue]) & ~ACTIVE) | BALANCED
    return np.array(out, np.int32).reshape(-1, 6).T.copy()


def raster(br, step1, step2):
    """raster_branches_cuda.cu:65-135: add every branch's jump to the step maps (int32,
    in place).  step1 (n1, n2, n3); step2 (n2, n3).  A branch to the border ends at the
    nearer border; one ending on the d1 = -1 side adds its jump to step2 of its column.
    Step1 positions are flat indices as the kernel computes them; one that falls outside
    the map (which the reference would write out of bounds) is dropped."""
    n1, n2, n3 = step1.shape
    if not (step1.flags.f_contiguous and step2.flags.f_contiguous):
        raise ValueError('step maps must be column-major (updated in place)')
    f1 = step1.reshape(-1, order='F')
    f2 = step2.reshape(-1, order='F')
    for b in range(br.shape[1]):
        p1x, p1y, p2x, p2y, z, jump = (int(v) for v in br[:, b])
        if p1x == -1:
            xs = p2x < n1 - p2x
            ys = p2y < n2 - p2y
            mdx = p2x if xs else n1 - p2x
            mdy = p2y if ys else n2 - p2y
            p1x = p2x if mdx > mdy else (-1 if xs else n1 - 1)
            p1y = p2y if mdx < mdy else (-1 if ys else n2 - 1)
            if p1x == -1:
                f2[n2 * z + p2y + 1] += jump
                continue
        if p1y != p2y:
            if p2y > p1y:
                p1x, p2x = p2x, p1x
                p1y, p2y = p2y, p1y
                jump = -jump
            steps = p1y - p2y
            for i in range(steps):
                cy = p2y + i + 1
                t = np.float32(np.float32(i) / np.float32(steps)) * np.float32(p1x - p2x)
                cx = p2x + int(np.float32(t)) + 1
                k = n1 * n2 * z + n1 * cy + cx
                if 0 <= k < f1.size:
                    f1[k] += jump
    return step1, step2


def step_maps(wrapped):
    """goldstein_branch_cut.m + remove_dipoles: the int32 step maps (step1 (n1, n2, n3),
    step2 (n2, n3)) and the intermediate products, for a wrapped (n1, n2, n3) phase."""
    w = np.asarray(wrapped, np.float32)
    res0 = residues(w)
    res, s1, s2 = remove_dipoles(res0)
    step1 = np.asfortranarray(np.round(s1).astype(np.int32))
    step2 = np.asfortranarray(np.round(s2).astype(np.int32))
    m1, m2, n3 = res.shape
    size3 = (m1, m2, n3)
    ordered, square = order_residues(res)
    if ordered.shape[1] == 0:
        return step1, step2, dict(residues=res0, dipole_free=res, ordered=ordered, branches=np.zeros((6, 0), np.int32))
    table, zstart = lookup(ordered, square)
    short = shortest_list(ordered, table, zstart, square, size3)
    br = branches(ordered, short, square, table, zstart, size3)
    raster(br, step1, step2)
    return step1, step2, dict(residues=res0, dipole_free=res, ordered=ordered, square=square, table=table,
                              zstart=zstart, shortest=short, branches=br)


def unwrap(wrapped):
    """unwrapp2_gpu.m:1-5 for (n1, n2, n3) float32 phase maps."""
    w = np.asarray(wrapped, np.float32)
    step1, step2, _ = step_maps(w)
    return raster_unwrap(w, step1, step2)
