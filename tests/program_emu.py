"""Pure-Python executor of the compiled LOD peel program (test helper).

It walks the program exported by sl_program_export with exactly the indexing
the GPU kernel uses (level task lists, compact mixed-radix tables, digit maps,
arena offsets, per-table exponents), so the program compiler and the kernel's
addressing can be checked against the oracle on a machine without a GPU.
"""
from __future__ import annotations

import math

import numpy as np

OPD = 12   # ints per OpDesc
PRD = 12   # ints per PrevDesc
CHD = 4    # ints per ChildDesc

AFF = 0x9C
KT = 0x6C


def kid_trait(mat, pat, i, j):
    ma = (AFF >> (2 * mat + i)) & 1
    pa = (AFF >> (2 * pat + j)) & 1
    return (KT >> (2 * ((ma << 1) | pa))) & 3


class Program:
    def __init__(self, words: np.ndarray):
        w = [int(x) for x in words]
        self.n_ops, self.n_levels, self.last_op, self.arena, npr, nch, nt = w[:7]
        p = 7
        self.ops = [w[p + OPD * i: p + OPD * (i + 1)] for i in range(self.n_ops)]
        p += OPD * self.n_ops
        self.prevs = [w[p + PRD * i: p + PRD * (i + 1)] for i in range(npr)]
        p += PRD * npr
        self.children = [w[p + CHD * i: p + CHD * (i + 1)] for i in range(nch)]
        p += CHD * nch
        self.tasks = [x & 0xFFFFFFFF for x in w[p: p + nt]]
        p += nt
        self.level_base = w[p: p + self.n_levels + 1]
        assert len(self.level_base) == self.n_levels + 1

    def check_bounds(self, npos):
        """Every table slot the kernel touches lies inside the arena."""
        for o, d in enumerate(self.ops):
            typ, k, peel, n_prev, tab_off, size = d[:6]
            assert tab_off >= 0 and tab_off + size * npos <= self.arena, (o, d)
            for q in range(n_prev):
                pr = self.prevs[d[10] + q]
                assert pr[1] + pr[2] * npos <= self.arena
        for L in range(self.n_levels):
            for t in range(self.level_base[L], self.level_base[L + 1]):
                op = self.tasks[t] >> 16
                d = self.ops[op]
                assert 0 <= t - d[11] < d[5]

    def run(self, case, dg, locus, fp32=False):
        """Returns ln L (with exponents) per position; dg=None gives P(T)."""
        npos = case.nlod
        N, F = case.N, case.F
        ftype = np.float32 if fp32 else float
        arena = [ftype(0)] * self.arena
        maxv = np.zeros((self.n_ops, npos), np.float32)
        trait = case.disease_prob.astype(np.float32) if fp32 else case.disease_prob
        code = [0] * N
        if dg is not None:
            col0 = dg[locus]
            col1 = dg[locus + 1]
            for p in range(F, N):
                code[p] = int(col0[2 * p]) | (int(col1[2 * p]) << 1) | (int(col0[2 * p + 1]) << 2) | (int(col1[2 * p + 1]) << 3)
        pt = case.partial_theta[locus] if dg is not None else 0.0

        def scale(op, pos):
            if not fp32 or maxv[op, pos] == 0:
                return 1.0, 0
            e = math.frexp(float(maxv[op, pos]))[1]  # max = m * 2^e, m in [0.5,1)
            return ftype(2.0 ** -e), e

        def prev_value(pr, ext, pos):
            idx = 0
            for j in range(pr[3]):
                wd = pr[4 + j] & 0xFFFFFFFF
                src, rb, stride = wd & 15, (wd >> 4) & 255, wd >> 12
                v = (ext >> (2 * src)) & 3
                idx += ((rb >> (2 * v)) & 3) * stride
            x = arena[pr[1] + pos * pr[2] + idx]
            if fp32:
                x = ftype(x * scale(pr[0], pos)[0])
            return x

        def rho(cd, i, j, th):
            t1, a1, t2, a2 = th
            r = ftype(1.0)
            r = ftype(r * (a1 if (cd & 1) == i else t1))
            r = ftype(r * (a2 if ((cd >> 1) & 1) == i else t2))
            r = ftype(r * (a1 if ((cd >> 2) & 1) == j else t1))
            r = ftype(r * (a2 if ((cd >> 3) & 1) == j else t2))
            return r

        q = ftype(0.25)
        for L in range(self.n_levels):
            base, end = self.level_base[L], self.level_base[L + 1]
            for pos in range(npos):
                tl, tr = pt * (pos + 1), pt * (npos - pos)
                th = (ftype(tl), ftype(1.0 - tl), ftype(tr), ftype(1.0 - tr))
                for t in range(base, end):
                    op = self.tasks[t] >> 16
                    cell = self.tasks[t] & 0xFFFF
                    typ, k, peel, n_prev, tab_off, size, md, pd, n_child, child_off, prev_off, task_base = self.ops[op]
                    prs = self.prevs[prev_off: prev_off + n_prev]
                    total = ftype(0.0)
                    if typ == 1:
                        mat, pat = (cell >> (2 * md)) & 3, (cell >> (2 * pd)) & 3
                        for ij in range(4):
                            i, j = ij >> 1, ij & 1
                            kt = kid_trait(mat, pat, i, j)
                            tmp = ftype(trait[peel, kt])
                            if tmp == 0:
                                continue
                            ext = cell | (kt << (2 * k))
                            for pr in prs:
                                tmp = ftype(tmp * prev_value(pr, ext, pos))
                            if dg is None:
                                tmp = ftype(tmp * q)
                            else:
                                tmp = ftype(tmp * ftype(q * rho(code[peel], i, j, th)))
                            total = ftype(total + tmp)
                    elif typ == 2:
                        cds = self.children[child_off: child_off + n_child]
                        for v in range(4):
                            tmp = ftype(trait[peel, v])
                            if tmp == 0:
                                continue
                            ext = cell | (v << (2 * k))
                            for pr in prs:
                                tmp = ftype(tmp * prev_value(pr, ext, pos))
                            cp = ftype(1.0)
                            for person, cdig, odig, pim in cds:
                                kid = (cell >> (2 * cdig)) & 3
                                other = (cell >> (2 * odig)) & 3
                                mat, pat = (v, other) if pim else (other, v)
                                ct = ftype(0.0)
                                for ij in range(4):
                                    i, j = ij >> 1, ij & 1
                                    if kid_trait(mat, pat, i, j) != kid:
                                        continue
                                    ct = ftype(ct + (q if dg is None else ftype(q * rho(code[person], i, j, th))))
                                cp = ftype(cp * ct)
                            tmp = ftype(tmp * cp)
                            total = ftype(total + tmp)
                    else:
                        for v in range(4):
                            tmp = ftype(trait[peel, v])
                            if tmp == 0:
                                continue
                            ext = cell | (v << (2 * k))
                            for pr in prs:
                                tmp = ftype(tmp * prev_value(pr, ext, pos))
                            total = ftype(total + tmp)
                    arena[tab_off + pos * size + (t - task_base)] = total
                    if fp32:
                        maxv[op, pos] = max(maxv[op, pos], np.float32(total))
        out = []
        last = self.ops[self.last_op]
        for pos in range(npos):
            Lv = float(arena[last[4] + pos * last[5]])
            e = sum(scale(o, pos)[1] for o in range(self.n_ops) if o != self.last_op) if fp32 else 0
            out.append(math.log(Lv) + e * math.log(2.0))
        return np.array(out)
