"""numpy interpreter of the kernel programs (TEST INFRASTRUCTURE ONLY).

Executes exactly what the HIP kernels are specified to do with the words the
engine compiles -- stage programs (tqsv_stage.hip), per-gate launches
(tqsv_gates.hip / tqsv_adjoint.hip) -- so the host-side planner, encoder and
matrix pool can be checked against the oracle on a machine without a GPU.
"""
from __future__ import annotations

import numpy as np

from torchquantum_amd import engine as E

R = E.R_BITS
NS = 1 << R


def _bits(x, nb):
    return [(x >> i) & 1 for i in range(nb)]


def _pc(x):
    return bin(int(x)).count("1")


def _op_mat(mats, w, row):
    off = w[1] + ((row * w[2]) if (w[0] & 256) else 0)
    return mats[off:]


def layout_offsets(words, which):
    """(thread offsets [2^T], slot offsets [NS]) of a stage's load ("ld") or
    store ("st") layout: the register local bits are the H_FLAGS mask
    (0 = canonical: the top R local bits), slot bit k is the k-th of them
    ascending, thread-index bit i the i-th other local bit."""
    h = words
    m, T = int(h[E.H_M]), int(h[E.H_T])
    L = [int(v) for v in h[E.H_POS:E.H_POS + m]]     # in-row position of each local bit
    shift = E.FL_LOAD_SHIFT if which == "ld" else E.FL_STORE_SHIFT
    mask = (int(h[E.H_FLAGS]) >> shift) & 0x1FFF
    regs = list(range(T, m)) if mask == 0 else [b for b in range(m) if (mask >> b) & 1]
    thr = [b for b in range(m) if b not in regs]
    tid = np.arange(1 << T)
    to = np.zeros(1 << T, dtype=np.int64)
    for i, b in enumerate(thr):
        to |= ((tid >> i) & 1) << L[b]
    ro = np.array([sum(((j >> k) & 1) << L[regs[k]] for k in range(R)) for j in range(NS)], dtype=np.int64)
    return to, ro


def stage_apply(words, st, mats, n, init_zero=False, adjoint=False, lam=None, grad=None, skip=(), no_store=False):
    """One stage over every group.  Adjoint: skip = program positions whose
    un-apply is left out (their gradient partials still count); no_store =
    psi and lam are not written back (the pruned kernels of stagejit)."""
    h = words
    n_ops, m, T, nnl = int(h[E.H_NOPS]), int(h[E.H_M]), int(h[E.H_T]), int(h[E.H_NNL])
    nlpos = [int(v) for v in h[E.H_POS + m:E.H_POS + n]]
    ops = h[int(h[E.H_OFF_OPS]):]
    gslot = h[int(h[E.H_OFF_GSLOT]):int(h[E.H_OFF_GSLOT]) + int(h[E.H_NGP])]
    tid = np.arange(1 << T)
    # the forward reads in the load layout and writes in the store layout;
    # the adjoint walks the program backwards, so the other way round
    toff_in, roff_in = layout_offsets(h, "st" if adjoint else "ld")
    toff_out, roff_out = layout_offsets(h, "ld" if adjoint else "st")
    bsz = st.shape[0]
    for row in range(bsz):
        for rest in range(1 << nnl):
            goff = sum(((rest >> i) & 1) << p for i, p in enumerate(nlpos))
            idx = goff + toff_in[:, None] + roff_in[None, :]
            odx = goff + toff_out[:, None] + roff_out[None, :]
            if init_zero:
                a = (idx == 0).astype(st.dtype)
            else:
                a = st[row, idx].copy()
            if adjoint:
                l = lam[row, idx].copy()
                gacc = np.zeros(len(gslot))
            order = range(n_ops) if not adjoint else range(n_ops - 1, -1, -1)
            for o in order:
                w = [int(v) for v in ops[8 * o:8 * o + 8]]
                typ = w[0] & 15
                if typ == E.OP_REMAP:
                    rec = h[w[7]:]
                    fr = [int(v) for v in rec[:T + R]]
                    to = [int(v) for v in rec[T + R:2 * (T + R)]]
                    if adjoint:
                        fr, to = to, fr
                    a = _remap(a, fr, to, T, m)
                    if adjoint:
                        l = _remap(l, fr, to, T, m)
                    continue
                M = _op_mat(mats, w, row)
                _check_flags(typ, w, M)
                # the stage's feature bits select the kernel instantiation: a
                # missing bit would make the kernel skip the op
                if typ == E.OP_MAT2:
                    assert int(h[E.H_FLAGS]) & E.FEAT_MAT2, "MAT2 op in a stage without the MAT2 feature"
                if typ == E.OP_MATK:
                    assert int(h[E.H_FLAGS]) & E.FEAT_MATK, "MATK op in a stage without the MATK feature"
                tok = (tid & w[4]) == w[5]
                rm, rv = w[6] & 255, (w[6] >> 8) & 255
                P = ((4 if (w[0] & E.F_MG) else (w[0] >> 4) & 15)) if w[7] >= 0 else 0
                if not adjoint:
                    a = _fwd(typ, a, M, w, tok, rm, rv, tid)
                else:
                    a2, l2, g = _bwd(typ, a, l, M, w, tok, rm, rv, tid, P)
                    if o not in skip:
                        a, l = a2, l2
                    for q in range(P):
                        gacc[w[7] + q] += g[q]
            if not no_store:
                st[row, odx] = a
            if adjoint:
                if not no_store:
                    lam[row, odx] = l
                for t, s in enumerate(gslot):
                    grad[row, s] += gacc[t]


def _check_flags(typ, w, M):
    """Fast-path flags must agree with the data the kernel would read."""
    fl = w[0]
    if fl & E.F_XPERM:
        assert typ == E.OP_MAT1 and np.allclose(M[:4], [0, 1, 1, 0]), "XPERM flag on a non-X matrix"
    if fl & E.F_REAL:
        assert typ == E.OP_MAT1 and np.allclose(np.imag(M[:4]), 0), "REAL flag on a complex matrix"
    if fl & E.F_NOCTRL:
        assert w[4] == 0 and (w[6] & 255) == 0, "NOCTRL flag with controls"
    if fl & E.F_SIGN:
        assert typ == E.OP_DIAG and np.isclose(M[0], -1) and w[3] == 0 and ((w[6] >> 16) & 255) == 0


def _remap(a, fr, to, T, m):
    nthr = a.shape[0]
    lds = np.zeros(1 << m, dtype=a.dtype)
    tid = np.arange(nthr)

    def lidx(rec):
        lt = np.zeros(nthr, dtype=np.int64)
        for i in range(T):
            lt |= np.where((tid >> i) & 1, rec[i], 0)
        lr = np.array([sum(rec[T + k] if (j >> k) & 1 else 0 for k in range(R)) for j in range(NS)])
        return lt[:, None] | lr[None, :]

    lds[lidx(fr)] = a
    return lds[lidx(to)]


def _pairs(K):
    return [(j, j | (1 << K)) for j in range(NS) if not (j >> K) & 1]


def _matk_slots(w3):
    if w3 == 4:
        regs = [0, 1, 2, 3]
    else:
        regs = [k for k in range(R) if k != w3]
    kb = len(regs)
    bases = [0] if kb == 4 else [0, 1 << w3]
    out = []
    for b in bases:
        out.append([b | sum((1 << regs[q]) for q in range(kb) if (r >> q) & 1) for r in range(1 << kb)])
    return out


def _fwd(typ, a, M, w, tok, rm, rv, tid):
    a = a.copy()
    if typ == E.OP_MAT1:
        K = w[3]
        for j, j1 in _pairs(K):
            ok = tok & ((j & rm) == rv)
            x0, x1 = a[:, j].copy(), a[:, j1].copy()
            a[:, j] = np.where(ok, M[0] * x0 + M[1] * x1, x0)
            a[:, j1] = np.where(ok, M[2] * x0 + M[3] * x1, x1)
    elif typ == E.OP_MAT2:
        kh, kl = w[3] & 15, (w[3] >> 4) & 15
        U = np.array(M[:16]).reshape(4, 4)
        for j in range(NS):
            if j & ((1 << kh) | (1 << kl)):
                continue
            idx = [j, j | (1 << kl), j | (1 << kh), j | (1 << kh) | (1 << kl)]
            ok = tok & ((j & rm) == rv)
            x = a[:, idx].copy()
            y = x @ U.T
            a[:, idx] = np.where(ok[:, None], y, x)
    elif typ == E.OP_DIAG:
        tpar = np.array([_pc(t & w[3]) for t in tid])
        rpm = (w[6] >> 16) & 255
        for j in range(NS):
            ok = tok & ((j & rm) == rv)
            par = (tpar + _pc(j & rpm)) & 1
            f = np.where(par == 1, M[1], M[0])
            a[:, j] = np.where(ok, f * a[:, j], a[:, j])
    elif typ == E.OP_MATK:
        for slots in _matk_slots(w[3]):
            D = len(slots)
            U = np.array(M[:D * D]).reshape(D, D)
            x = a[:, slots].copy()
            a[:, slots] = np.where(tok[:, None], x @ U.T, x)
    return a


def _bwd(typ, p, l, M, w, tok, rm, rv, tid, P):
    p, l = p.copy(), l.copy()
    g = np.zeros(8)
    if typ == E.OP_MAT1 and (w[0] & E.F_MG):
        # four sums S = Im <lam|s|x> (s = I, X, Y, Z) over x = psi before the un-apply
        K = w[3]
        U = np.array(M[:4]).reshape(2, 2)
        Ud = U.conj().T
        paulis = [np.eye(2), np.array([[0, 1], [1, 0]]), np.array([[0, -1j], [1j, 0]]), np.diag([1, -1])]
        for j, j1 in _pairs(K):
            ok = tok & ((j & rm) == rv)
            x = np.stack([p[:, j], p[:, j1]], -1)
            lm = np.stack([l[:, j], l[:, j1]], -1)
            for q, sg in enumerate(paulis):
                g[q] += np.imag(np.sum(np.conj(lm[ok]) * (x[ok] @ sg.T)))
            y = x @ Ud.T
            ln = lm @ Ud.T
            p[:, j] = np.where(ok, y[:, 0], x[:, 0])
            p[:, j1] = np.where(ok, y[:, 1], x[:, 1])
            l[:, j] = np.where(ok, ln[:, 0], lm[:, 0])
            l[:, j1] = np.where(ok, ln[:, 1], lm[:, 1])
        return p, l, g
    if typ == E.OP_MAT1:
        K = w[3]
        U = np.array(M[:4]).reshape(2, 2)
        Ud = U.conj().T
        for j, j1 in _pairs(K):
            ok = tok & ((j & rm) == rv)
            x = np.stack([p[:, j], p[:, j1]], -1)
            lm = np.stack([l[:, j], l[:, j1]], -1)
            y = x @ Ud.T
            for q in range(P):
                D = np.array(M[4 * (q + 1):4 * (q + 2)]).reshape(2, 2)
                v = y @ D.T
                g[q] += np.sum(np.where(ok, np.real(np.sum(np.conj(lm) * v, -1)), 0.0))
            ln = lm @ Ud.T
            p[:, j] = np.where(ok, y[:, 0], x[:, 0])
            p[:, j1] = np.where(ok, y[:, 1], x[:, 1])
            l[:, j] = np.where(ok, ln[:, 0], lm[:, 0])
            l[:, j1] = np.where(ok, ln[:, 1], lm[:, 1])
    elif typ == E.OP_MAT2:
        kh, kl = w[3] & 15, (w[3] >> 4) & 15
        U = np.array(M[:16]).reshape(4, 4)
        Ud = U.conj().T
        for j in range(NS):
            if j & ((1 << kh) | (1 << kl)):
                continue
            idx = [j, j | (1 << kl), j | (1 << kh), j | (1 << kh) | (1 << kl)]
            ok = tok & ((j & rm) == rv)
            x, lm = p[:, idx].copy(), l[:, idx].copy()
            y = x @ Ud.T
            for q in range(P):
                D = np.array(M[16 * (q + 1):16 * (q + 2)]).reshape(4, 4)
                v = y @ D.T
                g[q] += np.sum(np.where(ok, np.real(np.sum(np.conj(lm) * v, -1)), 0.0))
            p[:, idx] = np.where(ok[:, None], y, x)
            l[:, idx] = np.where(ok[:, None], lm @ Ud.T, lm)
    elif typ == E.OP_DIAG:
        tpar = np.array([_pc(t & w[3]) for t in tid])
        rpm = (w[6] >> 16) & 255
        for j in range(NS):
            ok = tok & ((j & rm) == rv)
            par = (tpar + _pc(j & rpm)) & 1
            f = np.where(par == 1, M[1], M[0])
            y = np.conj(f) * p[:, j]
            for q in range(P):
                df = np.where(par == 1, M[2 * (q + 1) + 1], M[2 * (q + 1)])
                g[q] += np.sum(np.where(ok, np.real(np.conj(l[:, j]) * df * y), 0.0))
            p[:, j] = np.where(ok, y, p[:, j])
            l[:, j] = np.where(ok, np.conj(f) * l[:, j], l[:, j])
    elif typ == E.OP_MATK:
        for slots in _matk_slots(w[3]):
            D = len(slots)
            U = np.array(M[:D * D]).reshape(D, D)
            Ud = U.conj().T
            x, lm = p[:, slots].copy(), l[:, slots].copy()
            p[:, slots] = np.where(tok[:, None], x @ Ud.T, x)
            l[:, slots] = np.where(tok[:, None], lm @ Ud.T, lm)
    return p, l, g


# ---------------------------------------------------------------- per-gate launches
def solo_apply(k: E.KOp, st, M, n, bstride, adjoint_of=False):
    """Semantics of tqsv_apply_{1q,2q,diag,swap,kq} on st [bsz, 2^n] (in place)."""
    bsz = st.shape[0]
    N = 1 << n
    j = np.arange(N)
    cm, cv = k.cmask, k.cval
    ok = (j & cm) == cv
    for row in range(bsz):
        m = M[row * bstride:] if bstride else M
        s = st[row]
        if k.kind == "mat1":
            t = 1 << k.bits[0]
            sel = ok & ((j & t) == 0)
            i0 = j[sel]
            i1 = i0 | t
            x0, x1 = s[i0].copy(), s[i1].copy()
            s[i0] = m[0] * x0 + m[1] * x1
            s[i1] = m[2] * x0 + m[3] * x1
        elif k.kind == "mat2":
            hi, lo = 1 << k.bits[0], 1 << k.bits[1]
            sel = ok & ((j & (hi | lo)) == 0)
            b = j[sel]
            idx = np.stack([b, b | lo, b | hi, b | hi | lo], -1)
            U = np.array(m[:16]).reshape(4, 4)
            s[idx] = s[idx] @ U.T
        elif k.kind == "swap":
            a, bb = 1 << k.bits[0], 1 << k.bits[1]
            sel = ok & ((j & (a | bb)) == 0)
            b = j[sel]
            x, y = s[b | bb].copy(), s[b | a].copy()
            s[b | bb], s[b | a] = y, x
        elif k.kind == "diag":
            pm = sum(1 << b for b in k.bits)
            par = np.array([_pc(x & pm) & 1 for x in j])
            f = np.where(par == 1, m[1], m[0])
            s[ok] = (f * s)[ok]
        else:
            kb = len(k.bits)
            offs = np.array([sum(1 << k.bits[q] for q in range(kb) if (r >> (kb - 1 - q)) & 1)
                             for r in range(1 << kb)])
            tm = sum(1 << b for b in k.bits)
            sel = ok & ((j & tm) == 0)
            b = j[sel]
            idx = b[:, None] + offs[None, :]
            U = np.array(m[:(1 << kb) ** 2]).reshape(1 << kb, 1 << kb)
            s[idx] = s[idx] @ U.T


def decode_gate_record(w):
    """A tqsv_run_gates record (16 int64 words) -> (KOp, matrix offset, row stride),
    read the way the C executor reads it."""
    kinds = {1: "mat1", 2: "mat2", 3: "swap", 4: "diag", 5: "matk"}
    kb = int(w[1])
    cm, cv = int(w[12]), int(w[13])
    cbits = tuple(b for b in range(64) if (cm >> b) & 1)
    k = E.KOp(kinds[int(w[0])], tuple(int(b) for b in w[2:2 + kb]), cbits, tuple((cv >> b) & 1 for b in cbits))
    return k, int(w[14]), int(w[15])


def run_compiled(c: E.Compiled, pool: np.ndarray, state_in, init_zero=False):
    """Forward of a compiled plan; mirrors engine.forward."""
    n, bsz = c.n, c.bsz
    st = np.zeros((bsz, 1 << n), dtype=np.complex128) if state_in is None else np.array(state_in, np.complex128)
    first = True
    for kind, payload in c.fwd_steps:
        if kind == "stage":
            base, stg = payload
            stage_apply(c.prog_host[base:], st, pool, n, init_zero=init_zero and first)
        else:
            if init_zero and first:
                st[:] = 0
                st[:, 0] = 1
            for w in payload:
                k, off, bs = decode_gate_record(w)
                solo_apply(k, st, pool[off:], n, bs)
        first = False
    if init_zero and first:
        st[:] = 0
        st[:, 0] = 1
    return st


def adjoint_compiled(c: E.Compiled, pool: np.ndarray, psi, lam, need_input_grad=True):
    """Reverse sweep; mirrors engine.adjoint (without need_input_grad: the
    pruned sweep of Compiled.prune_info, as the pruned compiled kernels run
    it).  Returns grads [bsz, n_slots]."""
    n, bsz = c.n, c.bsz
    grad = np.zeros((bsz, c.grad_width))
    steps = list(enumerate(c.steps))
    stop, skip = 0, {}
    if not need_input_grad:
        stop, skip = c.prune_info()
    for si, (kind, payload) in reversed(steps[stop:]):
        if kind == "stage":
            base, stg = payload
            stage_apply(c.prog_host[base:], psi, pool, n, adjoint=True, lam=lam, grad=grad,
                        skip=frozenset(skip.get(si, ())), no_store=not need_input_grad and si == stop)
            continue
        k = c.kops[payload]
        if k.P > 0:
            off, bs = c.data_off[payload]
            # U^dag on psi, derivative inner products, U^dag on lam -- the per-gate adjoint kernels
            _solo_adjoint_grad(k, psi, lam, pool[off:], bs, n, grad, c.scatter_col.get(payload, k.grad_slots[0]))
        else:
            if k.kind == "swap":
                solo_apply(k, psi, pool, n, 0)
                solo_apply(k, lam, pool, n, 0)
            else:
                off, bs = c.adj_off[payload]
                solo_apply(k, psi, pool[off:], n, bs)
                solo_apply(k, lam, pool[off:], n, bs)
    # Hermitian-generator sums -> parameter gradients (semantics of tqsv_mg_to_grad)
    for it in c.mg_items:
        for row in range(bsz):
            sI, sX, sY, sZ = grad[row, it["mg_slot"]:it["mg_slot"] + 4]
            U = pool[it["u_off"] + row * it["bstride"]:][:4].reshape(2, 2)
            for q in range(it["P"]):
                D = pool[it["u_off"] + row * it["bstride"] + 4 * (q + 1):][:4].reshape(2, 2)
                H = 1j * D @ U.conj().T
                a, d = np.real(H[0, 0] + H[1, 1]) / 2, np.real(H[0, 0] - H[1, 1]) / 2
                b, cy = np.real(H[0, 1] + H[1, 0]) / 2, np.imag(H[1, 0] - H[0, 1]) / 2
                grad[row, it["slot"][q]] += a * sI + b * sX + cy * sY + d * sZ
    # scatter columns of per-gate ops with non-consecutive slots
    for i, col, sl in c.scatter:
        for q, slot in enumerate(sl):
            grad[:, slot] += grad[:, col + q]
    return grad[:, :max(c.n_slots, 1)]


def _solo_adjoint_grad(k, psi, lam, M, bs, n, grad, g0):
    """The per-gate adjoint step; parameter q's gradient goes to column g0 + q."""
    bsz = psi.shape[0]
    d = 2 if k.kind == "diag" else k.dim() ** 2
    for row in range(bsz):
        m = M[row * bs:] if bs else M
        U = np.array(m[:d])
        Ds = [np.array(m[d * (q + 1):d * (q + 2)]) for q in range(k.P)]
        # build full operators on this row by applying the per-gate semantics to basis vectors
        ps = psi[row:row + 1].copy()
        ls = lam[row:row + 1].copy()
        y = _apply_adj(k, ps, U, n)
        for q, D in enumerate(Ds):
            v = y.copy()
            _apply_plain(k, v, D, n, zero_inactive=True)
            grad[row, g0 + q] += np.real(np.sum(np.conj(ls) * v))
        psi[row] = y[0]
        lam[row] = _apply_adj(k, ls, U, n)[0]


def _apply_plain(k, st, m, n, zero_inactive=False):
    N = 1 << n
    j = np.arange(N)
    ok = (j & k.cmask) == k.cval
    out = st.copy()
    solo_apply(k, out, m, n, 0)
    if zero_inactive:
        out[:, ~ok] = 0
    st[:] = out


def _apply_adj(k, st, U, n):
    if k.kind == "diag":
        Ud = np.conj(U)
    else:
        dd = int(round(np.sqrt(len(U))))
        Ud = np.conj(U.reshape(dd, dd).T).reshape(-1)
    out = st.copy()
    solo_apply(k, out, Ud, n, 0)
    return out
