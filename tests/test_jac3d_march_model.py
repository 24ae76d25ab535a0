"""CPU model of the first-order 3D Euler Jacobian march
(k_jac_fo_euler3d in csrc/pda_box_kernels.hip).

The model restates the kernel's index arithmetic (workgroup -> segment,
line, strip of planes; lane -> axis, cell, state offsets; the strip's plane
loop with the next plane's prefetch; the LDS block and diagonal slots with the momentum
exchange of the lane's axis; the copy-out chunks) and checks, for grids from
one box cell upwards, segments that end partial, and strip heights 1..64,
that
  - every state read (the current line's and the prefetched one) is a cell
    inside the grid, and each face a lane evaluates joins the two cells the
    blocks it feeds need: the -a block of a cell takes JL of the face below
    it along a, the +a block JR of the face above, the diagonal JR of the
    face below and JL of the face above;
  - every stage entry of the line's cells is written exactly once by a block
    store or by the x part of the diagonal, and the diagonal's y and z parts
    add to entries the x part wrote;
  - every entry lands at (p(r), p(c)) of the block slot the sorted CSR
    columns give its neighbour ([-z, -y, -x, self, +x, +y, +z]);
  - the copy-out chunks cover the line's values [0, ncell * 175) exactly
    once, read only written stage entries, start 16-byte aligned, and the
    CSR range they write is the segment's cells' values (rowbase of a real
    first-order pattern, boundary cells with fewer blocks);
  - every box cell is covered by exactly one (workgroup, line, lane).
Plain Python over small grids; no GPU.
"""
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "pressio-demoapps_amd", "csrc", "pda_box_kernels.hip")
E = 21
M, RS, PER = 5, 35, 175


def test_model_matches_kernel_source():
    txt = open(SRC).read()
    assert re.search(r"constexpr int J3SEG = (\d+);", txt).group(1) == str(E)
    body = txt[txt.index("k_jac_fo_euler3d(View P"):]
    body = body[:body.index("}  // namespace")]
    for line in [
        "const int ax = lane < E ? 0 : (lane < 2 * E ? 1 : 2);",
        "const int l = lane - ax * E;",
        "const int64_t seg = bid % nseg, jj = B.lo[1] + (bid / nseg) % bh;",
        "const int64_t strip = bid / (nseg * bh);",
        "const int64_t k1 = k0 + rows_per_strip < B.hi[2] ? k0 + rows_per_strip : B.hi[2];",
        "for (int64_t kk = k0; kk < k1; ++kk) {",
        "const int ci = own ? l : ncell - 1;",
        "const int64_t st = ax == 0 ? 1 : (ax == 1 ? nx : nxy);",
        "double *d1 = stage + (own ? l : 0) * PER + (2 - ax) * M;  // block -a",
        "double *d2 = stage + (own ? l : 0) * PER + (4 + ax) * M;  // block +a",
        "double *dd = stage + (own ? l : 0) * PER + 3 * M;         // diagonal",
        "load_cell<M>(U, c - st, qa);",
        "load_cell<M>(U, c + st, qc);",
        "const int64_t kn = kk + 1 < k1 ? kk + 1 : kk;",
        "const int64_t c = kn * nxy + jj * nx + iseg + ci;",
        "d1[pr(r) * RS + pr(c)] = h * Ph::entry<0>(P, A0, r, c);",
        "d2[pr(r) * RS + pr(c)] = -h * Ph::entry<0>(P, B1, r, c);",
        "part[r * M + c] = (Ph::entry<0>(P, A1, r, c) - Ph::entry<0>(P, B0, r, c)) * h;",
        "const int npairs = (total - head) >> 1;",
        "const int q = q0 < npairs ? q0 : npairs - 1;",
        "for (int it = 0; it < (E * PER / 2 + 63) / 64; ++it) {",
    ]:
        assert line in body, line


def pr(r, ax):
    return 1 + ax if r == 1 else (1 if r == 1 + ax else r)


def rowbase_of(n):
    """First CSR value of each cell of a non-periodic first-order 3D pattern:
    a cell's 5 rows hold 5 values per block, one block per in-grid
    neighbour plus the cell."""
    nx, ny, nz = n
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    nb = 1 + (i > 0) + (i < nx - 1) + (j > 0) + (j < ny - 1) + (k > 0) + (k < nz - 1)
    per = (M * M * nb).ravel()
    return np.concatenate([[0], np.cumsum(per)[:-1]]), per


def run_model(n, rps, bug=None):
    nx, ny, nz = n
    nxy = nx * ny
    lo, hi = (1, 1, 1), (nx - 1, ny - 1, nz - 1)
    bw, bh, bd = hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]
    nseg = (bw + E - 1) // E
    nstrip = (bd + rps - 1) // rps
    rowbase, per = rowbase_of(n)
    covered = np.zeros(nx * ny * nz, np.int64)

    def inside(ax, xi, jj, kk, o):
        """the cell o steps along axis ax from column xi of line (jj, kk)"""
        p = [xi, jj, kk]
        p[ax] += o
        return 0 <= p[0] < nx and 0 <= p[1] < ny and 0 <= p[2] < nz

    for bid in range(nseg * bh * nstrip):
        seg, j = bid % nseg, lo[1] + (bid // nseg) % bh
        strip = bid // (nseg * bh)
        iseg = lo[0] + seg * E
        ncell = min(bw - seg * E, E)
        k0 = lo[2] + strip * rps
        k1 = min(k0 + rps, hi[2])
        lanes = []
        for lane in range(64):
            ax = 0 if lane < E else (1 if lane < 2 * E else 2)
            l = lane - ax * E
            own = l < ncell
            ci = l if own else ncell - 1
            if bug == "unclamped":   # lanes past the last cell keep their own column
                ci = l
            xl = ax == 0
            st = (1, nx, nxy)[ax]
            sa, sc = st, st
            v1 = v2 = own
            c2 = l
            s2 = 4 + ax
            if bug == "slot":        # the +a block in the -a slot of the other side
                s2 = 4 - ax
            lanes.append((ax, l, xl, ci, sa, sc, own, v1, v2, c2, s2))
        for kk in range(k0, k1):
            for kq in (kk, kk + 1 if kk + 1 < k1 else kk):   # the plane and its prefetch
                for (ax, l, xl, ci, sa, sc, *_rest) in lanes:
                    for o in (-1, 0, 1):
                        assert inside(ax, iseg + ci, j, kq, o), (n, rps, bid, kq, l, ax, o)
            line0 = kk * nxy + j * nx + iseg
            # faces each lane evaluates: A = (c - sa, c), B = (c, c + sc)
            faces = []
            for (ax, l, xl, ci, sa, sc, *_r) in lanes:
                c = line0 + ci
                faces.append(((c - sa, c), (c, c + sc)))
            stride = (1, nx, nxy)
            # stage writes: index -> list of (kind, lane)
            writes = {}
            for lane, (ax, l, xl, ci, sa, sc, own, v1, v2, c2, s2) in enumerate(lanes):
                A, Bf = faces[lane]
                if v1:
                    cell = line0 + l
                    assert A == (cell - stride[ax], cell)     # JL of the face below
                    for r in range(M):
                        for c in range(M):
                            idx = l * PER + (2 - ax) * M + pr(r, ax) * RS + pr(c, ax)
                            writes.setdefault(idx, []).append(("blk", lane, l, 2 - ax))
                if v2:
                    cell = line0 + c2
                    assert Bf == (cell, cell + stride[ax])    # JR of the face above
                    for r in range(M):
                        for c in range(M):
                            idx = c2 * PER + s2 * M + pr(r, ax) * RS + pr(c, ax)
                            writes.setdefault(idx, []).append(("blk", lane, c2, 4 + ax))
                if own:
                    cell = line0 + l
                    assert A == (cell - stride[ax], cell)     # JR of the face below
                    assert Bf == (cell, cell + stride[ax])    # JL of the face above
            # diagonal: x stores, then y adds, then z adds
            diag_state = {}
            for phase in range(3):
                for lane, (ax, l, xl, ci, sa, sc, own, v1, v2, c2, s2) in enumerate(lanes):
                    if not own or ax != phase:
                        continue
                    for r in range(M):
                        for c in range(M):
                            rr, cc = (r, c) if phase == 0 else (pr(r, ax), pr(c, ax))
                            idx = l * PER + 3 * M + rr * RS + cc
                            if phase == 0:
                                assert idx not in diag_state
                                writes.setdefault(idx, []).append(("diag", lane, l, 3))
                                diag_state[idx] = 1
                            else:
                                assert diag_state.get(idx) == phase, (idx, phase)
                                diag_state[idx] = phase + 1
            assert all(v == 3 for v in diag_state.values())
            total = ncell * PER
            assert sorted(writes) == list(range(total)), (n, rps, bid, j)
            for idx, w in writes.items():
                assert len(w) == 1, (idx, w)
                _, _, cell, slot = w[0]
                assert idx // PER == cell and (idx % PER) % RS // M == slot
            # copy-out
            rb0 = int(rowbase[line0])
            for q in range(ncell):
                assert rowbase[line0 + q] == rb0 + q * PER and per[line0 + q] == PER
            head = rb0 & 1
            npairs = (total - head) >> 1
            assert (rb0 + head) % 2 == 0
            hits = np.zeros(total, np.int64)
            for it in range((E * PER // 2 + 63) // 64):
                for lane in range(64):
                    q = min(lane + 64 * it, npairs - 1)
                    for e in (head + 2 * q, head + 2 * q + 1):
                        assert 0 <= e < total
                        hits[e] += 1
            if head:
                hits[0] += 1
            if (total - head) & 1:
                hits[total - 1] += 1
            assert (hits >= 1).all()
            for (ax, l, xl, ci, sa, sc, own, *_r) in lanes:
                if own and xl:
                    covered[line0 + l] += 1
    i = np.arange(nx * ny * nz)
    ci, cj, ck = i % nx, (i // nx) % ny, i // nxy
    inbox = (ci >= lo[0]) & (ci < hi[0]) & (cj >= lo[1]) & (cj < hi[1]) & (ck >= lo[2]) & (ck < hi[2])
    assert (covered[inbox] == 1).all() and (covered[~inbox] == 0).all()


@pytest.mark.parametrize("n", [(3, 3, 3), (10, 9, 8), (8, 8, 8), (40, 7, 6), (45, 5, 4), (24, 4, 3), (66, 3, 5)])
@pytest.mark.parametrize("rps", [1, 2, 3, 7, 64])
def test_march_model(n, rps):
    run_model(n, rps)


@pytest.mark.parametrize("bug", ["unclamped", "slot"])
def test_model_flags_planted_mistakes(bug):
    """The model rejects lanes past the last cell of a partial segment
    reading their own column (out of the grid at the last segment), and a
    +a block stored in the slot of another neighbour."""
    with pytest.raises(AssertionError):
        run_model((19, 4, 3), 2, bug=bug)
