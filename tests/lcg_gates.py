"""Input generators of the reference's own acceptance gates, restated.

Each generator replays the fixed-seed 64-bit LCG sweep of one reference test, draw for draw, and
yields the cases it builds.  Expected outputs come from the oracle; the point of replaying the
exact sweeps is that they are the inputs the reference's authors chose to gate a backend on.

  ext_single_cases   crates/bwa-extend/src/lib.rs:283-354   assert_backend_matches_scalar
  ext_batch_rounds   crates/bwa-extend/src/lib.rs:426-522   assert_backend_batch_matches_scalar
  ext_read_sized     crates/bwa-neon/src/lib.rs:276-372     neon_batch_matches_scalar_read_sized
  ext_ungapped       crates/bwa-neon/src/lib.rs:374-470     neon_ungapped_hit_matches_scalar
  rescue_jobs        crates/bwa-neon/src/matesw.rs:5045-5139 matesw_equals_scalar
  unbanded_cases     crates/bwa-extend/src/sw.rs:1326-1356  matches_unbanded_reference
"""

MASK = (1 << 64) - 1
LCG_MUL = 6364136223846793005
LCG_INC = 1442695040888963407


class Lcg:
    def __init__(self, seed):
        self.s = seed & MASK

    def __call__(self):
        self.s = (self.s * LCG_MUL + LCG_INC) & MASK
        return self.s >> 33


def scmat(a, b):
    m = []
    for i in range(4):
        for j in range(4):
            m.append(a if i == j else -b)
        m.append(-1)
    m += [-1] * 5
    return m


def ext_single_cases():
    nxt = Lcg(0x9E3779B97F4A7C15)
    for case in range(2000):
        qlen = 1 + nxt() % 60
        tlen = 1 + nxt() % 60
        q = [nxt() % 4 for _ in range(qlen)]
        t = [nxt() % 4 for _ in range(tlen)]
        w = 1 + nxt() % 30
        zdrop = nxt() % 120
        end_bonus = nxt() % 10
        h0 = 1 + nxt() % 20
        if case == 0:
            a, b, o_del, e_del, o_ins, e_ins = 1, 4, 6, 1, 6, 1
        else:
            a = 1 + nxt() % 4
            b = 1 + nxt() % 6
            o_del = 1 + nxt() % 8
            e_del = 1 + nxt() % 3
            o_ins = 1 + nxt() % 8
            e_ins = 1 + nxt() % 3
        yield dict(q=q, t=t, mat=scmat(a, b), o_del=o_del, e_del=e_del, o_ins=o_ins,
                   e_ins=e_ins, w=w, end_bonus=end_bonus, zdrop=zdrop, h0=h0)


def ext_batch_rounds():
    nxt = Lcg(0x123456789ABCDEF0)
    sizes = [1, 2, 3, 4, 7, 8, 9, 15, 16, 17, 20, 31]
    for _ in range(200):
        bs = sizes[nxt() % 12]
        w = 1 + nxt() % 30
        zdrop = nxt() % 120
        end_bonus = nxt() % 10
        o_del = 4 + nxt() % 4
        e_del = 1 + nxt() % 2
        o_ins = 4 + nxt() % 4
        e_ins = 1 + nxt() % 2
        qs = []
        for _ in range(bs):
            ql = 1 + nxt() % 80
            qs.append([nxt() % 4 for _ in range(ql)])
        ts = []
        for _ in range(bs):
            tl = 1 + nxt() % 80
            ts.append([nxt() % 4 for _ in range(tl)])
        h0s = [1 + nxt() % 20 for _ in range(bs)]
        yield dict(queries=qs, targets=ts, h0s=h0s, mat=scmat(1, 4), o_del=o_del, e_del=e_del,
                   o_ins=o_ins, e_ins=e_ins, w=w, end_bonus=end_bonus, zdrop=zdrop)


def ext_read_sized():
    nxt = Lcg(0xC0FFEE123456789A)
    sizes = [1, 5, 8, 9, 16, 17, 24, 40]
    for _ in range(120):
        bs = sizes[nxt() % 8]
        w = 50 + nxt() % 120
        zdrop = nxt() % 200
        end_bonus = nxt() % 12
        o_del = 4 + nxt() % 4
        e_del = 1 + nxt() % 2
        o_ins = 4 + nxt() % 4
        e_ins = 1 + nxt() % 2
        qs, ts, h0s = [], [], []
        for _ in range(bs):
            ql = 60 + nxt() % 200
            q = [nxt() % 4 for _ in range(ql)]
            tl = ql + nxt() % 60
            t = []
            qi = 0
            while len(t) < tl:
                if qi < len(q) and nxt() % 100 >= 4:
                    t.append(q[qi])
                    qi += 1
                else:
                    t.append(nxt() % 4)
                    if nxt() % 2 == 0:
                        qi += 1
            qs.append(q)
            ts.append(t)
            h0s.append(20 + nxt() % 30)
        yield dict(queries=qs, targets=ts, h0s=h0s, mat=scmat(1, 4), o_del=o_del, e_del=e_del,
                   o_ins=o_ins, e_ins=e_ins, w=w, end_bonus=end_bonus, zdrop=zdrop)


def ext_ungapped():
    nxt = Lcg(0x51ED5EEDD1A60000)
    sizes = [1, 4, 8, 13, 16, 24]
    for _ in range(300):
        bs = sizes[nxt() % 6]
        w = 1 + nxt() % 150
        zdrop = nxt() % 200
        end_bonus = nxt() % 12
        qs, ts, h0s = [], [], []
        for _ in range(bs):
            ql = 1 + nxt() % 200
            q = [nxt() % 4 for _ in range(ql)]
            mode = nxt() % 5
            extra = nxt() % 40
            t = list(q)
            if mode == 0:
                t += [nxt() % 4 for _ in range(extra)]
            elif mode == 1:
                pass
            elif mode == 2:
                cut = max(0, ql - (1 + nxt() % 3))
                t = t[:cut]
                t += [nxt() % 4 for _ in range(extra)]
            elif mode == 3:
                for _ in range(1 + nxt() % 2):
                    if t:
                        p = nxt() % len(t)
                        t[p] = (t[p] + 1 + nxt() % 3) % 4
                t += [nxt() % 4 for _ in range(extra)]
            else:
                if t:
                    p = nxt() % len(t)
                    t[p] = 4
                t += [nxt() % 4 for _ in range(extra)]
            qs.append(q)
            ts.append(t)
            h0s.append(1 + nxt() % 40)
        yield dict(queries=qs, targets=ts, h0s=h0s, mat=scmat(1, 4), o_del=6, e_del=1, o_ins=6,
                   e_ins=1, w=w, end_bonus=end_bonus, zdrop=zdrop)


def rescue_jobs():
    """The 2000 jobs of matesw_equals_scalar; run under scorings (1,4,minsc 19) and (10,40,190)."""
    nxt = Lcg(0x123456789ABCDEF1)
    qs, ts = [], []
    for _ in range(2000):
        ql = 5 + nxt() % 146
        tl = ql + nxt() % 500
        t = [nxt() % 5 for _ in range(tl)]
        q = [nxt() % 5 for _ in range(ql)]
        copies = 1 + nxt() % 2
        if nxt() % 5 != 0:
            for _ in range(copies):
                if tl > ql:
                    at = nxt() % (tl - ql + 1)
                    t[at:at + ql] = q
            for _ in range(nxt() % 4):
                p = nxt() % ql
                q[p] = nxt() % 4
        if nxt() % 4 == 0:
            q[nxt() % ql] = 4
        if nxt() % 4 == 0:
            t[nxt() % tl] = 4
        qs.append(q)
        ts.append(t)
    return qs, ts


RESCUE_SCORINGS = ((1, 4, 19), (10, 40, 190))


def unbanded_cases():
    """matches_unbanded_reference: xorshift64 stream, 500 (query, target) pairs."""
    state = [0xA5A5123499990001]

    def nxt():
        s = state[0]
        s ^= (s << 13) & MASK
        s ^= s >> 7
        s ^= (s << 17) & MASK
        state[0] = s
        return s

    for _ in range(500):
        ln = 20 + nxt() % 40
        base = [nxt() % 4 for _ in range(ln)]
        target = list(base)
        if ln > 25:
            p = ln - 3 - nxt() % 3
            target[p] = (target[p] + 1) % 4
        yield base, target
