"""Forward-mode (dual-number) derivative of the 1D2V Strang splitting.

TEST INFRASTRUCTURE: a checker of the reverse mode, never the product.

The reference is Julia code written for Float64 (particle_mesh_coupling_1d.jl:
261-266 types its arguments), so its dual-number derivative cannot be taken
by running it; this module restates the splitting's operators on dual numbers
a + b eps (eps^2 = 0), vectorised over particles with numpy, and carries a
tangent direction through strang_splitting! (hamiltonian_splitting.jl:98-108)
with the loss the reverse mode differentiates,

    J = w * sum_{s=1..T} 1/2 l2norm_squared(B_s, p-1) + <ct, final state>.

The forward map follows the oracle (gempic_oracle.c, the reference's
operators):
  operatorHB   compute_e_from_b!(E2, dt, B)                 (1d2v.jl:234-236)
  operatorHE   v += dt qm E(x); compute_b_from_e!(B, dt, E2)       (:191-219)
  operatorHp2  v1 += dt qm v2 B(x); j2 = dt sum q v2 S0(x); E2 -= M0^{-1} j2/dx
                                                                   (:129-176)
  operatorHp1  x += dt v1 with add_current_update_v! (trunc() cells, the old,
               new and in-between segments, Gauss-Legendre on each), v2 update,
               x = mod(x, L); E1 -= M1^{-1} j1 / dx                  (:41-112)
The field operators are linear in the fields, so their tangent is the same
operator applied to the tangent; they come from the oracle (pyoracle.Maxwell1D).
Each particle quantity is an independent computation (this file does not call
the oracle's particle functions), so agreement with the oracle's reverse mode
checks both.
"""
from __future__ import annotations

import numpy as np

import pyoracle as orc


class Dual:
    """Arrays a + b eps with eps^2 = 0."""

    __slots__ = ("a", "b")

    def __init__(self, a, b=None):
        self.a = np.asarray(a, dtype=np.float64)
        self.b = np.zeros_like(self.a) if b is None else np.asarray(b, dtype=np.float64)

    @staticmethod
    def _c(o):
        return o if isinstance(o, Dual) else Dual(o, 0.0 * np.asarray(o, dtype=np.float64))

    def __add__(self, o):
        o = Dual._c(o)
        return Dual(self.a + o.a, self.b + o.b)

    __radd__ = __add__

    def __sub__(self, o):
        o = Dual._c(o)
        return Dual(self.a - o.a, self.b - o.b)

    def __rsub__(self, o):
        return Dual._c(o) - self

    def __mul__(self, o):
        o = Dual._c(o)
        return Dual(self.a * o.a, self.a * o.b + self.b * o.a)

    __rmul__ = __mul__

    def __neg__(self):
        return Dual(-self.a, -self.b)

    def __getitem__(self, k):
        return Dual(self.a[k], self.b[k])

    def __setitem__(self, k, v):
        v = Dual._c(v)
        self.a[k] = v.a
        self.b[k] = v.b

    def copy(self):
        return Dual(self.a.copy(), self.b.copy())


def bspline(degree, off):
    """uniform_bsplines_eval_basis! (low_level_bsplines.jl:63-80) on duals."""
    b = [Dual(np.ones_like(off.a))] + [None] * degree
    for j in range(1, degree + 1):
        xx = -off
        jr = float(j)
        inv_j = 1.0 / jr
        saved = Dual(np.zeros_like(off.a))
        for r in range(j):
            xx = xx + 1.0
            temp = b[r] * inv_j
            b[r] = saved + xx * temp
            saved = (jr - xx) * temp
        b[j] = saved
    return b


class Tangent1D2V:
    """The 1D2V splitting of nx cells on [xmin, xmax) with smoothers of
    degrees (p0, p1) (galerkin) and a Maxwell1DFEM of degree maxwell_degree
    (the loss's l2norm_squared takes its s_deg_1)."""

    def __init__(self, nx, xmin, xmax, p0, p1, maxwell_degree=3, charge=1.0, mass=1.0,
                 common_weight=1.0):
        self.nx, self.xmin, self.L = nx, xmin, xmax - xmin
        self.dx = self.L / nx
        self.p0, self.p1 = p0, p1
        self.charge, self.qm, self.cw = charge, charge / mass, common_weight
        self.m = orc.Maxwell1D(xmin, xmax, nx, maxwell_degree)
        self.s_deg_1 = maxwell_degree - 1
        self.nq = (p1 + 2) // 2
        self.qx, self.qw = orc.gausslegendre(self.nq)

    # -- linear field maps: the same operator on the value and on the tangent
    def _apply(self, fn, target, *args):
        """fn(target_array, *args) in place, linear in (target, args)."""
        ta = target.a.copy()
        fn(ta, *[a.a for a in args])
        tb = target.b.copy()
        fn(tb, *[a.b for a in args])
        return Dual(ta, tb)

    # -- particle kernels
    def _cell(self, x):
        xi = Dual((x.a - self.xmin) / self.dx, x.b / self.dx)
        idx = np.trunc(xi.a).astype(np.int64)
        off = xi - idx.astype(np.float64)
        return idx, off

    def evaluate(self, degree, x, f):
        idx, off = self._cell(x)
        b = bspline(degree, off)
        v = Dual(np.zeros_like(x.a))
        for k in range(degree + 1):
            node = np.mod(idx - degree + k, self.nx)
            v = v + f[node] * b[k]
        return v

    def deposit(self, degree, x, q):
        idx, off = self._cell(x)
        b = bspline(degree, off)
        j = Dual(np.zeros(self.nx))
        for k in range(degree + 1):
            node = np.mod(idx - degree + k, self.nx)
            c = q * b[k]
            np.add.at(j.a, node, c.a)
            np.add.at(j.b, node, c.b)
        return j

    def _update_jv(self, j, lower, upper, index, q, sign, v, bfield, mask):
        """update_jv! (particle_mesh_coupling_1d.jl:385-414) for the particles in mask."""
        if not np.any(mask):
            return v
        lo, up, ix, qq = lower[mask], upper[mask], index[mask], q[mask]
        c1 = (up - lo) * 0.5
        c2 = (up + lo) * 0.5
        sv = None
        for iq in range(self.nq):
            b = bspline(self.p1, c1 * self.qx[iq] + c2)
            part = [bk * (self.qw[iq] * 1.0) * c1 for bk in b]
            sv = part if sv is None else [s + pp for s, pp in zip(sv, part)]
        vm = v[mask]
        for k in range(self.p1 + 1):
            val = sv[k] * (sign * self.dx)
            node = np.mod(ix - self.p1 + k, self.nx)
            c = qq * val
            np.add.at(j.a, node, c.a)
            np.add.at(j.b, node, c.b)
            vm = vm - self.qm * val * bfield[node]
        v[mask] = vm
        return v

    def current_update_v(self, x_old, x_new, q, bfield, v):
        """add_current_update_v! with B (particle_mesh_coupling_1d.jl:296-376)."""
        j = Dual(np.zeros(self.nx))
        io, ro = self._cell(x_old)
        inw, rn = self._cell(x_new)
        one = Dual(np.ones_like(ro.a))
        zero = Dual(np.zeros_like(ro.a))
        same = io == inw
        fwd = io < inw
        bwd = io > inw
        lt = ro.a < rn.a
        v = self._update_jv(j, ro, rn, io, q, 1.0, v, bfield, same & lt)
        v = self._update_jv(j, rn, ro, io, q, -1.0, v, bfield, same & ~lt)
        v = self._update_jv(j, ro, one, io, q, 1.0, v, bfield, fwd)
        v = self._update_jv(j, zero, rn, inw, q, 1.0, v, bfield, fwd)
        v = self._update_jv(j, rn, one, inw, q, -1.0, v, bfield, bwd)
        v = self._update_jv(j, zero, ro, io, q, -1.0, v, bfield, bwd)
        gap = np.abs(inw - io) - 1
        for k in range(1, int(gap.max(initial=0)) + 1):
            mf = fwd & (gap >= k)
            v = self._update_jv(j, zero, one, io + k, q, 1.0, v, bfield, mf)
            mb = bwd & (gap >= k)
            v = self._update_jv(j, zero, one, inw + k, q, -1.0, v, bfield, mb)
        return j, v

    # -- operators on the state (x, v1, v2, w as arrays; E1, E2, B as Duals)
    def hb(self, st, dt):
        st["e2"] = self._apply(lambda e, b: self.m.compute_e_from_b(e, dt, b), st["e2"], st["b"])

    def he(self, st, dt):
        x = st["x"]
        e1 = self.evaluate(self.p1, x, st["e1"])
        e2 = self.evaluate(self.p0, x, st["e2"])
        st["v1"] = st["v1"] + dt * self.qm * e1
        st["v2"] = st["v2"] + dt * self.qm * e2
        st["b"] = self._apply(lambda b, e: self.m.compute_b_from_e(b, dt, e), st["b"], st["e2"])

    def hp2(self, st, dt):
        x, v2 = st["x"], st["v2"]
        bb = self.evaluate(self.p1, x, st["b"])
        st["v1"] = st["v1"] + dt * self.qm * v2 * bb
        q = st["w"] * self.charge * self.cw * v2
        j2 = self.deposit(self.p0, x, q) * dt
        st["e2"] = self._apply(lambda e, j: self.m.compute_e_from_j(e, np.ascontiguousarray(j), 2),
                               st["e2"], j2)

    def hp1(self, st, dt):
        x, v1 = st["x"], st["v1"]
        x_new = x + dt * v1
        q = st["w"] * self.charge * self.cw
        j1, v2 = self.current_update_v(x, x_new, q, st["b"], st["v2"].copy())
        st["v2"] = v2
        # mod(x, L): the tangent passes through (the shift is a constant)
        st["x"] = Dual(self.xmin + np.mod(x_new.a - self.xmin, self.L), x_new.b)
        st["e1"] = self._apply(lambda e, j: self.m.compute_e_from_j(e, np.ascontiguousarray(j), 1),
                               st["e1"], j1)

    def strang_jvp(self, parts, e1, e2, b, d_parts, d_e1, d_e2, d_b, dt, nsteps, wloss,
                   ct_parts=None, ct_e1=None, ct_e2=None, ct_b=None):
        """(J, dJ) at the state (parts rows x, v1, v2, w; e1, e2, b) in the
        direction (d_parts rows x, v1, v2; d_e1, d_e2, d_b)."""
        st = {"x": Dual(parts[0], d_parts[0]), "v1": Dual(parts[1], d_parts[1]),
              "v2": Dual(parts[2], d_parts[2]), "w": Dual(parts[3], np.zeros_like(parts[3])),
              "e1": Dual(e1, d_e1), "e2": Dual(e2, d_e2), "b": Dual(b, d_b)}
        J, dJ = 0.0, 0.0
        h = 0.5 * dt
        for _ in range(nsteps):
            self.hb(st, h)
            self.he(st, h)
            self.hp2(st, h)
            self.hp1(st, dt)
            self.hp2(st, h)
            self.he(st, h)
            self.hb(st, h)
            B = st["b"]
            J += wloss * 0.5 * self.m.l2norm_squared(np.ascontiguousarray(B.a), self.s_deg_1)
            dJ += wloss * self.m.inner_product(np.ascontiguousarray(B.a),
                                               np.ascontiguousarray(B.b), self.s_deg_1)
        for key, row, ct in (("x", 0, None), ("v1", 1, None), ("v2", 2, None)):
            if ct_parts is not None:
                J += float(np.dot(ct_parts[row], st[key].a))
                dJ += float(np.dot(ct_parts[row], st[key].b))
        for key, ct in (("e1", ct_e1), ("e2", ct_e2), ("b", ct_b)):
            if ct is not None:
                J += float(np.dot(ct, st[key].a))
                dJ += float(np.dot(ct, st[key].b))
        return J, dJ, st
