"""Zernike surface distortion (drop-in for raypier.core.cdistortions).

The reference evaluates radial polynomials by a memoised recursion over a
3 x k_max workspace (cdistortions.pyx:149-316).  Here the recursion is run
once, symbolically, when the scene is compiled: each value the reference
would compute becomes one instruction writing the same workspace slot, and
every cached read becomes a reference to that slot.  The GPU then executes
the straight-line program per point.  Slot sharing in the reference's
derivative recursion (its C-term index uses (h-1)*h + |m|, :211) is kept, so
cache hits resolve exactly as they do in the reference.
"""
import math

import numpy as np

from . import _abi
from .ctracer import Distortion

# ANSI j -> (n, m, k) table, grown on demand as in cdistortions.pyx:101-136
jnm_map = [(0, 0, 0)]


def eval_nmk(j):
    """(n, m, k) for ANSI index j (cdistortions.pyx:104-136)."""
    if j < 0:
        raise ValueError("J must be non-negative")
    if j >= len(jnm_map):
        n, m, k = jnm_map[-1]
        while True:
            while m < n:
                m += 2
                next_j = (n * (n + 2) + m) // 2
                half_n = int(math.floor(n / 2.))
                k = half_n * (half_n + 1) + abs(m)
                jnm_map.append((n, m, k))
                if next_j == j:
                    return (n, m, k)
                elif next_j > j:
                    raise ValueError("Something went wrong here! Missed j value.")
            n += 1
            m = -2 - n
    return jnm_map[j]


def _cdiv(a, b):
    """C integer division (truncation toward zero)."""
    q = abs(a) // abs(b)
    return q if (a >= 0) == (b > 0) else -q


CONST_ZERO = -1
CONST_ONE = -2


class _ZProgram(object):
    """Symbolic execution of the reference recursion over workspace slots."""

    def __init__(self, kmax, init_r00=False):
        self.kmax = kmax
        self.written = set()
        self.insns = []
        if init_r00:   # z_offset_c sets workspace[0,0] = 1.0 (cdistortions.pyx:431)
            self._emit(2, self._slot(0, 0), [CONST_ONE, CONST_ZERO, CONST_ZERO, CONST_ZERO, CONST_ZERO])

    def _slot(self, row, k):
        s = row * self.kmax + k
        if s < 0 or s >= 3 * self.kmax:
            raise ValueError("Zernike workspace index out of range")
        return s

    def _emit(self, op, dst, src):
        rec = np.zeros((), _abi.zernike_insn_dtype)
        rec["op"] = op
        rec["dst"] = dst
        rec["src"] = src + [CONST_ZERO] * (5 - len(src))
        self.insns.append(rec)
        self.written.add(dst)

    def R(self, k, n, m):   # zernike_R_c, :149-178
        if n < m:
            return CONST_ZERO
        if n == 0:
            return CONST_ONE
        s = self._slot(0, k)
        if s in self.written:
            return s
        nA, mA = n - 1, abs(m - 1)
        h = _cdiv(nA, 2)
        kA = h * (h + 1) + abs(mA)
        nB, mB = nA, m + 1
        kB = h * (h + 1) + abs(mB)
        nC, mC = n - 2, m
        h = _cdiv(nC, 2)
        kC = h * (h + 1) + abs(mC)
        a = self.R(kA, nA, mA)
        b = self.R(kB, nB, mB)
        c = self.R(kC, nC, mC)
        self._emit(0, s, [a, b, c])
        return s

    def Rp(self, k, n, m):   # zernike_Rprime_c, :189-219
        if n < m:
            return CONST_ZERO
        if n == 0:
            return CONST_ZERO
        s = self._slot(1, k)
        if s in self.written:
            return s
        nA, mA = n - 1, abs(m - 1)
        h = _cdiv(nA, 2)
        kA = h * (h + 1) + abs(mA)
        nB, mB = nA, m + 1
        kB = h * (h + 1) + abs(mB)
        nC, mC = n - 2, m
        kC = (h - 1) * h + abs(mC)
        a = self.R(kA, nA, mA)
        b = self.R(kB, nB, mB)
        c = self.Rp(kA, nA, mA)
        d = self.Rp(kB, nB, mB)
        e = self.Rp(kC, nC, mC)
        self._emit(1, s, [a, b, c, d, e])
        return s

    def Ror(self, k, n, m):   # zernike_R_over_r_c, :288-316
        if n < m:
            return CONST_ZERO
        s = self._slot(2, k)
        if s in self.written:
            return s
        nA, mA = n - 1, abs(m - 1)
        h = _cdiv(nA, 2)
        kA = h * (h + 1) + abs(mA)
        nB, mB = nA, m + 1
        kB = h * (h + 1) + abs(mB)
        nC, mC = n - 2, m
        h = _cdiv(nC, 2)
        kC = h * (h + 1) + abs(mC)
        a = self.R(kA, nA, mA)
        b = self.R(kB, nB, mB)
        c = self.Ror(kC, nC, mC)
        self._emit(2, s, [a, b, c])
        return s


def compact_slots(insns, final_refs):
    """Renumber the workspace slots of one straight-line program so that a
    slot is reused once its value is dead (linear scan over the program;
    the slots in final_refs stay live to the end).  Returns (program, map
    from old to new slot, slots used).  Every source is written before it
    is read (memoised recursion), so the renumbering keeps every value."""
    last_use = {}
    for k, ins in enumerate(insns):
        for s in ins["src"]:
            if s >= 0:
                last_use[int(s)] = k
    for s in final_refs:
        last_use[s] = len(insns)
    mapping, free, used = {}, [], 0
    out = []
    for k, ins in enumerate(insns):
        rec = ins.copy()
        rec["src"] = [mapping[int(s)] if s >= 0 else int(s) for s in ins["src"]]
        for s in sorted(set(int(x) for x in ins["src"] if x >= 0)):
            if last_use.get(s) == k:
                free.append(mapping[s])
        free.sort(reverse=True)
        dst = int(ins["dst"])
        if free:
            phys = free.pop()
        else:
            phys = used
            used += 1
        mapping[dst] = phys
        rec["dst"] = phys
        if last_use.get(dst, -1) < k:     # written and never read: free at once
            free.append(phys)
        out.append(rec)
    return out, mapping, used


def _padd(a, b):
    n = max(len(a), len(b))
    return [(a[k] if k < len(a) else 0) + (b[k] if k < len(b) else 0) for k in range(n)]


def _pneg(a):
    return [-x for x in a]


def _pshift(a):          # r * a
    return [0] + list(a)


def program_polynomials(insns):
    """Run a straight-line program on integer polynomials in r (lists of
    Python ints, ascending powers): every op is r*(a+b)-c, (a+b)+r*(c+d)-e or
    (a+b)-c, so each slot holds a polynomial with integer coefficients, the
    exact expansion of what the recursion computes.  Returns value(ref)."""
    slot = {}

    def val(ref):
        ref = int(ref)
        if ref == CONST_ZERO:
            return [0]
        if ref == CONST_ONE:
            return [1]
        return slot[ref]

    for ins in insns:
        a, b, c, d, e = (val(x) for x in ins["src"])
        op = int(ins["op"])
        if op == 0:
            v = _padd(_pshift(_padd(a, b)), _pneg(c))
        elif op == 1:
            v = _padd(_padd(_padd(a, b), _pshift(_padd(c, d))), _pneg(e))
        else:
            v = _padd(_padd(a, b), _pneg(c))
        while len(v) > 1 and v[-1] == 0:
            v.pop()
        slot[int(ins["dst"])] = v
    return val


def compile_zernike(coef_list, unit_radius):
    """Compile sorted (j, value) pairs into (zernike record fields, coef
    records, sag instructions, grad instructions)."""
    nmk = [eval_nmk(j) for j, _ in coef_list]
    n_max = max(n for n, m, k in nmk)
    kmax = (n_max // 2) * (n_max // 2 + 1) + n_max + 1   # cdistortions.pyx:391
    if 3 * kmax > _abi.MAX_ZERNIKE_SLOTS:
        raise ValueError("Zernike order too high for the device workspace (n <= 9)")
    sag = _ZProgram(kmax, init_r00=True)
    grad = _ZProgram(kmax)
    coefs = []
    for (j, v), (n, m, k) in zip(coef_list, nmk):
        rec = np.zeros((), _abi.zernike_coef_dtype)
        rec["n"], rec["m"], rec["value"] = n, m, v
        # N = sqrt(n+1) or sqrt(2(n+1)), then N *= value, as the reference
        # evaluates it per call (correctly rounded on host and device alike)
        N = math.sqrt(float(n + 1)) if m == 0 else math.sqrt(float(2 * (n + 1)))
        rec["nval"] = N * float(v)
        rec["sag_ref"] = sag.R(k, n, abs(m))
        # z_offset_and_gradient_c calls R, R', R/r in this order (:496-498)
        rec["r_ref"] = grad.R(k, n, abs(m))
        rec["rp_ref"] = grad.Rp(k, n, abs(m))
        rec["ror_ref"] = grad.Ror(k, n, abs(m))
        coefs.append(rec)
    # the values as polynomials in r (what the kernels evaluate)
    sag_val, grad_val = program_polynomials(sag.insns), program_polynomials(grad.insns)
    for c in coefs:
        for k, (f, val) in enumerate((("sag_ref", sag_val), ("r_ref", grad_val), ("rp_ref", grad_val),
                                      ("ror_ref", grad_val))):
            p = val(c[f])
            if len(p) > _abi.ZPOLY or any(abs(x) >= 2 ** 53 for x in p):
                raise ValueError("Zernike polynomial out of range")
            c["deg"][k] = len(p) - 1
            c["poly"][k][:len(p)] = [float(x) for x in p]
    sag_p, smap, s_used = compact_slots(sag.insns, [int(c["sag_ref"]) for c in coefs if c["sag_ref"] >= 0])
    grad_p, gmap, g_used = compact_slots(grad.insns, [int(c[f]) for c in coefs for f in ("r_ref", "rp_ref", "ror_ref")
                                                      if c[f] >= 0])
    for c in coefs:
        if c["sag_ref"] >= 0:
            c["sag_ref"] = smap[int(c["sag_ref"])]
        for f in ("r_ref", "rp_ref", "ror_ref"):
            if c[f] >= 0:
                c[f] = gmap[int(c[f])]
    return kmax, coefs, sag_p, grad_p, max(s_used, g_used, 1)


class ZernikeDistortion(Distortion):
    """Sum of Zernike terms over ANSI indices j (cdistortions.pyx:323-514).

    ZernikeDistortion(j7=1.0, unit_radius=10.0) or ZernikeDistortion([(j, v), ...])."""

    def __init__(self, *args, **coefs):
        self.coef_map = {}
        self.unit_radius = coefs.get("unit_radius", 1.0)
        cdict = {int(k[1:]): float(v) for k, v in coefs.items() if k.startswith("j")}
        if args:
            for k, v in args[0]:
                cdict[int(k)] = float(v)
        self.k_max = 0
        self.set_coefs(list(cdict.items()))

    def set_coefs(self, coefs):
        clist = sorted(coefs)
        self.j_max = max(j for j, v in clist)
        self.n_coefs = len(clist)
        self._coefs = clist
        self.coef_map = {j: i for i, (j, v) in enumerate(clist)}
        n_max = max(eval_nmk(j)[0] for j, v in clist)
        self.k_max = max(self.k_max, (n_max // 2) * (n_max // 2 + 1) + n_max + 1)

    def __getitem__(self, idx):
        if idx >= self.n_coefs:
            raise IndexError("Index %d greater than number of coeffs (%d)." % (idx, self.n_coefs))
        j, v = self._coefs[idx]
        n, m, k = eval_nmk(j)
        return (j, n, m, k, v)

    def __len__(self):
        return self.n_coefs

    def update_coef(self, j, value):
        if j not in self.coef_map:
            self.append_coef(j, value)
        i = self.coef_map[j]
        self._coefs[i] = (j, float(value))

    def get_coef(self, j):
        return self._coefs[self.coef_map[j]][1]

    def append_coef(self, k, value):
        raise NotImplementedError()

    def _rpx_compile(self, rec, builder):
        kmax, coefs, sag, grad, n_slots = compile_zernike(self._coefs, self.unit_radius)
        cb, ib = builder.add_zernike_program(coefs, sag + grad)
        rec["unit_radius"] = self.unit_radius
        rec["n_coefs"] = len(coefs)
        rec["coef_begin"] = cb
        rec["sag_begin"] = ib
        rec["sag_len"] = len(sag)
        rec["grad_begin"] = ib + len(sag)
        rec["grad_len"] = len(grad)
        rec["n_slots"] = n_slots

    def z_offset(self, x, y):
        from ._queries import zernike_eval
        return zernike_eval(self, x, y)[0]

    def z_offset_and_gradient(self, x, y):
        from ._queries import zernike_eval
        return zernike_eval(self, x, y)[1]
