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
SQD = 24   # ints per SeqOp: type, n_prev, size, out_off, idx_base, tinfo_base, peel, n_child,
           # prev_off[8], child[7], pad

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
        p += self.n_levels + 1
        self.waves = w[p]
        p += 1
        self.sched_off = w[p: p + self.n_levels * self.waves + 1]
        p += self.n_levels * self.waves + 1
        self.sched_ops = w[p: p + self.n_ops]
        p += self.n_ops
        nti = w[p]
        self.tinfo = [x & 0xFFFFFFFF for x in w[p + 1: p + 1 + nti]]
        p += 1 + nti
        nix = w[p]
        self.idx = w[p + 1: p + 1 + nix]
        p += 1 + nix
        self.seq_elems = w[p]
        p += 1
        self.seq = [w[p + SQD * i: p + SQD * (i + 1)] for i in range(self.n_ops)]
        p += SQD * self.n_ops
        nls = w[p]
        self.lstream = np.array(w[p + 1: p + 1 + nls], dtype=np.int64).astype(np.uint32)
        assert p + 1 + nls == len(w)

    def check_terms(self, case):
        """The precompiled term words and table indices agree with what the
        digit maps give (kid trait of (i, j), validity, child masks, indices)."""
        for o, d in enumerate(self.ops):
            typ, k, peel, n_prev, tab_off, size, md, idx_base, n_child, child_off, prev_off, task_base = d
            prs = self.prevs[prev_off: prev_off + n_prev]
            cds = self.children[child_off: child_off + n_child]
            for c in range(size):
                cell = self.tasks[task_base + c] & 0xFFFF
                for t in range(4):
                    if typ == 1:
                        mother, father = int(case.mother[peel]), int(case.father[peel])
                        cut = list(case.ops.cutset(o))
                        mat = (cell >> (2 * cut.index(mother))) & 3
                        pat = (cell >> (2 * cut.index(father))) & 3
                        val = kid_trait(mat, pat, t >> 1, t & 1)
                    else:
                        val = t
                    info = self.tinfo[(task_base + c) * 4 + t]
                    assert info & 3 == val
                    assert bool(info & 4) == (case.disease_prob[peel, val] != 0)
                    for ch, (person, cdig, odig, pim) in enumerate(cds):
                        kid = (cell >> (2 * cdig)) & 3
                        other = (cell >> (2 * odig)) & 3
                        mat, pat = (val, other) if pim else (other, val)
                        m = sum(1 << ij for ij in range(4) if kid_trait(mat, pat, ij >> 1, ij & 1) == kid)
                        assert (info >> (4 + 4 * ch)) & 15 == m
                    ext = cell | (val << (2 * k))
                    for q, pr in enumerate(prs):
                        if not info & 4:
                            assert self.idx[idx_base + (c * 4 + t) * n_prev + q] == 0
                            continue
                        ix = 0
                        for j in range(pr[3]):
                            wd = pr[4 + j] & 0xFFFFFFFF
                            v = (ext >> (2 * (wd & 15))) & 3
                            ix += ((((wd >> 4) & 255) >> (2 * v)) & 3) * (wd >> 12)
                        assert self.idx[idx_base + (c * 4 + t) * n_prev + q] == ix

    def check_schedule(self):
        """Every op is run exactly once, by one wave, in its own level."""
        seen = []
        for L in range(self.n_levels):
            level_ops = {self.tasks[t] >> 16 for t in range(self.level_base[L], self.level_base[L + 1])}
            ran = []
            for w in range(self.waves):
                a, b = self.sched_off[L * self.waves + w], self.sched_off[L * self.waves + w + 1]
                ran += self.sched_ops[a:b]
            assert sorted(ran) == sorted(level_ops), L
            seen += ran
        assert sorted(seen) == list(range(self.n_ops))

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
        maxv = np.zeros(self.n_ops, np.float32)
        trait = case.disease_prob.astype(np.float32) if fp32 else case.disease_prob
        code = [0] * N
        if dg is not None:
            col0 = dg[locus]
            col1 = dg[locus + 1]
            for p in range(F, N):
                code[p] = int(col0[2 * p]) | (int(col1[2 * p]) << 1) | (int(col0[2 * p + 1]) << 2) | (int(col1[2 * p + 1]) << 3)
        pt = case.partial_theta[locus] if dg is not None else 0.0

        def scale(op, pos):  # one exponent per table, shared by all positions
            if not fp32 or maxv[op] == 0:
                return 1.0, 0
            e = math.frexp(float(maxv[op]))[1]  # max = m * 2^e, m in [0.5,1)
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
                    typ, k, peel, n_prev, tab_off, size, md, idx_base, n_child, child_off, prev_off, task_base = self.ops[op]
                    c = t - task_base
                    prs = self.prevs[prev_off: prev_off + n_prev]
                    kids = [cd[0] for cd in self.children[child_off: child_off + n_child]]
                    total = ftype(0.0)
                    for term in range(4):
                        info = self.tinfo[t * 4 + term]
                        if not info & 4:
                            continue
                        tmp = ftype(trait[peel, info & 3])
                        for qq, pr in enumerate(prs):
                            ix = self.idx[idx_base + (c * 4 + term) * n_prev + qq]
                            x = arena[pr[1] + pos * pr[2] + ix]
                            if fp32:
                                x = ftype(x * scale(pr[0], pos)[0])
                            tmp = ftype(tmp * x)
                        if typ == 1:
                            if dg is None:
                                tmp = ftype(tmp * q)
                            else:
                                tmp = ftype(tmp * ftype(q * rho(code[peel], term >> 1, term & 1, th)))
                        elif typ == 2:
                            cp = ftype(1.0)
                            for ch, person in enumerate(kids):
                                m = (info >> (4 + 4 * ch)) & 15
                                ct = ftype(0.0)
                                for ij in range(4):
                                    if (m >> ij) & 1:
                                        ct = ftype(ct + (q if dg is None else ftype(q * rho(code[person], ij >> 1, ij & 1, th))))
                                cp = ftype(cp * ct)
                            tmp = ftype(tmp * cp)
                        total = ftype(total + tmp)
                    arena[tab_off + pos * size + c] = total
                    if fp32:
                        maxv[op] = max(maxv[op], np.float32(total))
        out = []
        last = self.ops[self.last_op]
        for pos in range(npos):
            Lv = float(arena[last[4] + pos * last[5]])
            e = sum(scale(o, pos)[1] for o in range(self.n_ops) if o != self.last_op) if fp32 else 0
            out.append(math.log(Lv) + e * math.log(2.0))
        return np.array(out)

    # ------------------------------------------------------------ sequential form
    def check_seq(self):
        """The sequential form runs every op once, after the tables it
        consumes, and no two tables alive at the same time share arena cells;
        every touched cell lies inside seq_elems."""
        by_op = {}
        for s in self.seq:
            typ, n_prev, size, out, idx_base, tinfo_base, peel, n_child = s[:8]
            by_op[tinfo_base // 4] = s
        live = {}  # out offset -> (size, op position)
        done_offsets = []
        for pos, s in enumerate(self.seq):
            typ, n_prev, size, out = s[:4]
            assert 0 <= out and out + size <= self.seq_elems
            consumed = s[8:8 + n_prev]
            for off in consumed:
                assert off in live, (pos, off)
            for off, (sz, _) in live.items():
                if off in consumed:
                    continue
                assert out + size <= off or off + sz <= out, ("overlap", pos)
            for off in consumed:
                # the op's output may not overlap what it reads
                sz = live[off][0]
                assert out + size <= off or off + sz <= out, ("reads its own output", pos)
            for off in consumed:
                del live[off]
            live[out] = (size, pos)
        assert len(live) == 1 and self.seq[-1][2] == 1

    def run_seq(self, case, dg, locus, fp32=False):
        """The lane kernel's arithmetic for one (graph, position) per lane:
        ops in sequential order, cells in order, FP32 tables renormalised by a
        power of two when their largest cell leaves [2^-6, 2^6).  Returns ln L
        per position."""
        npos = case.nlod
        N, F = case.N, case.F
        ftype = np.float32 if fp32 else float
        trait = case.disease_prob.astype(np.float32) if fp32 else case.disease_prob
        code = [0] * N
        col0, col1 = dg[locus], dg[locus + 1]
        for p in range(F, N):
            code[p] = int(col0[2 * p]) | (int(col1[2 * p]) << 1) | (int(col0[2 * p + 1]) << 2) | (int(col1[2 * p + 1]) << 3)
        pt = case.partial_theta[locus]
        q = ftype(0.25)
        out = []
        for pos in range(npos):
            tl, tr = pt * (pos + 1), pt * (npos - pos)
            t1, a1, t2, a2 = ftype(tl), ftype(1.0 - tl), ftype(tr), ftype(1.0 - tr)

            def qrho(cd, i, j):
                r = ftype(1.0)
                r = ftype(r * (a1 if (cd & 1) == i else t1))
                r = ftype(r * (a2 if ((cd >> 1) & 1) == i else t2))
                r = ftype(r * (a1 if ((cd >> 2) & 1) == j else t1))
                r = ftype(r * (a2 if ((cd >> 3) & 1) == j else t2))
                return ftype(q * r)

            arena = [ftype(0)] * self.seq_elems
            E = 0
            for s in self.seq:
                typ, n_prev, size, out_off, idx_base, tinfo_base, peel, n_child = s[:8]
                poff = s[8:8 + n_prev]
                kids = s[16:16 + n_child]
                qr = [qrho(code[peel], t >> 1, t & 1) for t in range(4)] if typ == 1 else None
                r4 = [[qrho(code[kk], ij >> 1, ij & 1) for ij in range(4)] for kk in kids]
                mx = ftype(0)
                for c in range(size):
                    total = ftype(0.0)
                    for term in range(4):
                        info = self.tinfo[tinfo_base + 4 * c + term]
                        if not info & 4:
                            continue
                        tmp = ftype(trait[peel, info & 3])
                        for qq in range(n_prev):
                            ix = self.idx[idx_base + (c * 4 + term) * n_prev + qq]
                            tmp = ftype(tmp * arena[poff[qq] + ix])
                        if typ == 1:
                            tmp = ftype(tmp * qr[term])
                        elif typ == 2:
                            cp = ftype(1.0)
                            for ch in range(n_child):
                                m = (info >> (4 + 4 * ch)) & 15
                                ct = ftype(0.0)
                                for ij in range(4):
                                    if (m >> ij) & 1:
                                        ct = ftype(ct + r4[ch][ij])
                                cp = ftype(cp * ct)
                            tmp = ftype(tmp * cp)
                        total = ftype(total + tmp)
                    arena[out_off + c] = total
                    mx = max(mx, total)
                if fp32 and mx > 0:
                    e = math.frexp(float(mx))[1] - 1  # mx in [2^e, 2^(e+1))
                    if e < -6 or e >= 6:
                        for c in range(size):
                            arena[out_off + c] = ftype(arena[out_off + c] * ftype(2.0 ** -e))
                        E += e
            L = float(arena[self.seq[-1][3]])
            out.append(math.log(L) + E * math.log(2.0))
        return np.array(out)

    # ------------------------------------------------------------ lane stream
    def run_lane_stream(self, case, dg, locus):
        """The FP32 lane kernel's execution of its descriptor stream for one
        graph at one interval: per position, the wavefront's transmission
        table Z[code][mask] / Q[code][t], the lane's 4-bit codes, the ops'
        headers and 32-word cell blocks (padding cells included), and the
        power-of-two renormalisation.  Returns ln L per position."""
        f32 = np.float32
        S = self.lstream
        assert S.size > 0
        N, F = case.N, case.F
        col0, col1 = dg[locus], dg[locus + 1]
        code = [int(col0[2 * p]) | (int(col1[2 * p]) << 1) | (int(col0[2 * p + 1]) << 2) | (int(col1[2 * p + 1]) << 3)
                for p in range(F, N)]
        pt = case.partial_theta[locus]
        npos = case.nlod
        out = []
        for pos in range(npos):
            tl, tq = pt * (pos + 1), pt * (npos - pos)
            t1, a1, t2, a2 = f32(tl), f32(1.0 - tl), f32(tq), f32(1.0 - tq)

            def zentry(cd, mask):
                a, b = f32(0), f32(0)
                for v in range(2):
                    f = f32(f32(0.25) * f32((a1 if (cd & 1) == v else t1) * (a2 if ((cd >> 1) & 1) == v else t2)))
                    g = f32((a1 if ((cd >> 2) & 1) == v else t1) * (a2 if ((cd >> 3) & 1) == v else t2))
                    if (mask >> v) & 1:
                        a = f32(a + f)
                    if (mask >> (2 + v)) & 1:
                        b = f32(b + g)
                return f32(a * b)

            Z = [[zentry(cd, m) for m in range(16)] for cd in range(16)]
            Q = [[Z[cd][(1 << (t >> 1)) | (4 << (t & 1))] for t in range(4)] for cd in range(16)]
            arena = np.zeros(self.seq_elems, f32)
            E, h, last = 0, 0, 0
            for _ in range(self.n_ops):
                hd = S[h:h + 32]
                typ, npv, size, outo, peel_r, nc, stride, nblk = (int(x) for x in hd[:8])
                tr = hd[16:20].view(np.float32)
                last = outo
                qr = Q[code[peel_r]] if typ == 1 else None
                kid = [code[int(hd[8 + ch])] for ch in range(nc)] if typ == 2 else []
                G, cpb = 1 + npv, 32 // stride
                assert stride == (8 if npv <= 1 else 16 if npv <= 3 else 32)
                mx = f32(0)
                for b in range(nblk):
                    for k in range(cpb):
                        c = b * cpb + k
                        rec = S[h + 32 + 32 * b + k * stride: h + 32 + 32 * b + (k + 1) * stride]
                        total = f32(0)
                        for t in range(4):
                            w = int(rec[t * G])
                            tmp = f32(rec[t * G: t * G + 1].view(np.float32)[0] * qr[t]) if typ == 1 else tr[t]
                            for q in range(npv):
                                off = int(rec[t * G + 1 + q])
                                assert off % 256 == 0 and off // 256 < self.seq_elems
                                tmp = f32(tmp * arena[off // 256])
                            for ch in range(nc if typ == 2 else 0):
                                tmp = f32(tmp * Z[kid[ch]][(w >> (4 * ch)) & 15])
                            total = f32(total + tmp)
                        assert outo + c < self.seq_elems
                        arena[outo + c] = total
                        if not np.isnan(total):
                            mx = max(mx, total)
                e = math.frexp(float(mx))[1] - 1 if mx > 0 else 0
                if -6 <= e < 6:
                    e = 0
                if e:
                    arena[outo:outo + size] = np.ldexp(arena[outo:outo + size], -e).astype(f32)
                    E += e
                h += 32 + 32 * nblk
            out.append(math.log(float(arena[last])) + E * math.log(2.0))
        return np.array(out)
