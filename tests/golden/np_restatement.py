"""Independent numpy restatement of rustpde's 2-D RBC step (test infrastructure).

This file is NOT the product and NOT the C++ oracle.  It is a second, independent
restatement of the reference algorithm, written with numpy/scipy (pocketfft DCT-I and
real FFT, LAPACK dgeev/getri through numpy.linalg), used only to

* generate the golden fixtures under tests/golden/ (make_golden.py), and
* cross-check the C++ oracle (oracle/) on the CPU.

Every routine cites the reference file:line it restates (paths relative to the
reference repository root).  Arrays are row-major [nx, ny] like ndarray.
"""
import numpy as np
from scipy.fft import dct

# ----------------------------------------------------------------------------------
# 1-D bases  (funspace/src/chebyshev/*.rs, funspace/src/fourier/r2c.rs)
# ----------------------------------------------------------------------------------


class Base:
    """kind in {'cheb','cd','cn','cdbc','fo'}; n = physical length."""

    def __init__(self, kind, n):
        self.kind, self.n = kind, n
        if kind == "cheb":
            self.m = n
        elif kind in ("cd", "cn"):
            self.m = n - 2
        elif kind == "cdbc":
            self.m = 2
        elif kind == "fo":
            self.m = n // 2 + 1
        else:
            raise ValueError(kind)
        self.ortho_len = self.m if kind == "fo" else n
        if kind in ("cd", "cn"):
            m = n - 2
            self.diag = np.ones(m)
            if kind == "cd":  # composite_stencil.rs:117-135
                self.low2 = -np.ones(m)
            else:  # composite_stencil.rs:144-162
                k = np.arange(m, dtype=float)
                self.low2 = -1.0 * (k ** 2) / ((k + 2) ** 2)
            # _get_main_off composite_stencil.rs:165-176
            self.main = self.diag * self.diag + self.low2 * self.low2
            self.off = self.diag[2:] * self.low2[:-2]
        if kind == "cdbc":  # composite_stencil.rs:287-296
            self.t0 = np.array([0.5, 0.5])
            self.t1 = np.array([-0.5, 0.5])

    @property
    def is_fourier(self):
        return self.kind == "fo"

    # nodes ortho.rs:71-80 ; c2c.rs:63-66
    def coords(self):
        n = self.n
        if self.kind == "fo":
            return np.arange(0.0, 2 * np.pi, 2 * np.pi / n)
        m = float(n - 1)
        k = np.arange(n, dtype=float)
        return -np.sin(np.pi * (m - 2.0 * k) / (2.0 * m))

    # stencil as matrix composite_stencil.rs:180-188, 319-326
    def stencil(self):
        n = self.n
        if self.kind in ("cd", "cn"):
            S = np.zeros((n, n - 2))
            for i in range(n - 2):
                S[i, i] = self.diag[i]
                S[i + 2, i] = self.low2[i]
            return S
        if self.kind == "cdbc":
            S = np.zeros((n, 2))
            S[0, 0], S[0, 1] = self.t0
            S[1, 0], S[1, 1] = self.t1
            return S
        raise ValueError

    # mass() ortho.rs:206-208, composite.rs:338-340, r2c.rs:115-117
    def mass(self):
        if self.kind == "cheb":
            return np.eye(self.n)
        if self.kind == "fo":
            return np.eye(self.m)
        return self.stencil()

    # laplace() ortho.rs:486-488 (dmsuite.rs:24-49), r2c.rs:372-378
    def laplace(self):
        if self.kind == "fo":
            k = np.arange(self.m, dtype=float)
            return np.diag(-k * k)
        n = self.n
        D = np.zeros((n, n))
        for p in range(n):
            for q in range(p + 2, n):
                if (p + q) % 2 == 0:
                    D[p, q] = float(q * (q * q - p * p))
        D[0, :] *= 0.5
        return D

    # pinv ortho.rs:147-174 ; r2c.rs:381-387
    def laplace_inv(self):
        if self.kind == "fo":
            P = self.laplace()
            for i in range(1, self.m):
                P[i, i] = 1.0 / P[i, i]
            return P
        n = self.n
        P = np.zeros((n, n))
        P[2, 0] = 0.25
        for i in range(3, n):
            P[i, i - 2] = 1.0 / float(4 * i * (i - 1))
        for i in range(2, n - 2):
            P[i, i] = -1.0 / float(2 * (i * i - 1))
        for i in range(2, n - 4):
            P[i, i + 2] = 1.0 / float(4 * i * (i + 1))
        return P

    # laplace_inv_eye ortho.rs:177-180 ; r2c.rs:391-394
    def laplace_inv_eye(self):
        if self.kind == "fo":
            return np.eye(self.m)[1:, :]
        return np.eye(self.n)[2:, :]


def _moveaxis_apply(a, axis, fn):
    b = np.moveaxis(a, axis, -1)
    r = fn(b)
    return np.moveaxis(r, -1, axis)


def cheb_forward(v):
    """ortho.rs:235-258: DCT-I, then x (-1)^k/(n-1), endpoints halved (last axis)."""
    n = v.shape[-1]
    y = dct(v, type=1, axis=-1)
    corr = np.array([(-1.0) ** i for i in range(n)]) * (1.0 / (n - 1))
    y = y * corr
    y[..., 0] *= 0.5
    y[..., n - 1] *= 0.5
    return y


def cheb_backward(c):
    """ortho.rs:288-312: x (-1)^k/2, endpoints doubled, then DCT-I (last axis)."""
    n = c.shape[-1]
    corr = np.array([(-1.0) ** i for i in range(n)]) / 2.0
    b = c * corr
    b[..., 0] *= 2.0
    b[..., n - 1] *= 2.0
    return dct(b, type=1, axis=-1)


def cheb_diff(c, n_times):
    """ortho.rs:107-125 differentiate_lane, vectorised over leading axes."""
    d = np.array(c, copy=True)
    n = d.shape[-1]
    for _ in range(n_times):
        d[..., 0] = d[..., 1]
        for i in range(1, n - 1):
            d[..., i] = 2.0 * float(i + 1) * d[..., i + 1]
        d[..., n - 1] = 0.0
        for i in range(n - 3, 0, -1):
            d[..., i] = d[..., i] + d[..., i + 2]
        d[..., 0] = d[..., 0] + d[..., 2] / 2.0
    return d


def tdma(a, b, c, d):
    """funspace/src/chebyshev/linalg.rs:14-57 (offsets -2,0,2), last axis of d."""
    n = d.shape[-1]
    w = np.zeros(n - 2)
    g = np.zeros(d.shape, dtype=d.dtype)
    w[0] = c[0] / b[0]
    g[..., 0] = d[..., 0] / b[0]
    if len(c) > 1:
        w[1] = c[1] / b[1]
    g[..., 1] = d[..., 1] / b[1]
    for i in range(2, n - 2):
        w[i] = c[i] / (b[i] - a[i - 2] * w[i - 2])
    for i in range(2, n):
        g[..., i] = (d[..., i] - g[..., i - 2] * a[i - 2]) / (b[i] - a[i - 2] * w[i - 2])
    x = np.zeros(d.shape, dtype=d.dtype)
    x[..., n - 1] = g[..., n - 1]
    x[..., n - 2] = g[..., n - 2]
    for i in range(n - 2, 0, -1):
        x[..., i - 1] = g[..., i - 1] - x[..., i + 1] * w[i - 1]
    return x


def base_to_ortho(base, c):
    """composite_stencil.rs:207-229 / :355-360 (last axis)."""
    if base.kind in ("cheb", "fo"):
        return np.array(c, copy=True)
    n = base.n
    p = np.zeros(c.shape[:-1] + (n,), dtype=c.dtype)
    if base.kind == "cdbc":
        p[..., 0] = c[..., 0] * base.t0[0] + c[..., 1] * base.t0[1]
        p[..., 1] = c[..., 0] * base.t1[0] + c[..., 1] * base.t1[1]
        return p
    p[..., 0] = c[..., 0] * base.diag[0]
    p[..., 1] = c[..., 1] * base.diag[1]
    for i in range(2, n - 2):
        p[..., i] = c[..., i] * base.diag[i] + c[..., i - 2] * base.low2[i - 2]
    p[..., n - 2] = c[..., n - 4] * base.low2[n - 4]
    p[..., n - 1] = c[..., n - 3] * base.low2[n - 3]
    return p


def base_from_ortho(base, p):
    """composite_stencil.rs:250-276 / :366-388 (last axis)."""
    if base.kind in ("cheb", "fo"):
        return np.array(p, copy=True)
    if base.kind == "cdbc":
        t0, t1 = base.t0, base.t1
        c0 = p[..., 0] * t0[0] + p[..., 1] * t1[0]
        c1 = p[..., 0] * t0[1] + p[..., 1] * t1[1]
        a = t0[0] * t0[0] + t1[0] * t1[0]
        b = t0[0] * t0[1] + t1[0] * t1[1]
        c = t0[1] * t0[0] + t1[1] * t1[0]
        d = t0[1] * t0[1] + t1[1] * t1[1]
        det = 1.0 / (a * d - b * c)
        out = np.zeros(p.shape[:-1] + (2,), dtype=p.dtype)
        out[..., 0] = (c0 * d - c1 * b) * det
        out[..., 1] = (c1 * a - c0 * c) * det
        return out
    m = base.m
    cc = p[..., :m] * base.diag + p[..., 2:m + 2] * base.low2
    return tdma(base.off, base.main, base.off, cc)


def base_forward(base, v):
    """last axis: ortho.rs:235-258, composite.rs:383-396, r2c.rs:160-175."""
    if base.kind == "fo":
        return np.fft.rfft(v, axis=-1)
    p = cheb_forward(v)
    return base_from_ortho(base, p)


def base_backward(base, c):
    """last axis: ortho.rs:288-312, composite.rs:430-443, r2c.rs:211-226 (1/n)."""
    if base.kind == "fo":
        return np.fft.irfft(c, n=base.n, axis=-1)
    return cheb_backward(base_to_ortho(base, c))


def base_differentiate(base, c, n_times):
    """ortho.rs:107-125, composite.rs:523-536, r2c.rs:88-99 (last axis)."""
    if base.kind == "fo":
        k = 1j * np.arange(base.m, dtype=float)
        d = np.array(c, copy=True)
        for _ in range(n_times):
            d = d * k
        return d
    return cheb_diff(base_to_ortho(base, c), n_times)


class Space2:
    """funspace/src/space2.rs:136-356 composition over two axes."""

    def __init__(self, b0, b1):
        self.b = (b0, b1)

    @property
    def is_complex(self):
        return self.b[0].is_fourier

    def shape_phys(self):
        return (self.b[0].n, self.b[1].n)

    def shape_spec(self):
        return (self.b[0].m, self.b[1].m)

    def _ax(self, base, a, axis, fn):
        return _moveaxis_apply(a, axis, lambda x: fn(base, x))

    def forward(self, v):  # axis 1 then axis 0
        buf = self._ax(self.b[1], v, 1, base_forward)
        return self._ax(self.b[0], buf, 0, base_forward)

    def backward(self, vhat):  # axis 0 then axis 1
        buf = self._ax(self.b[0], vhat, 0, base_backward)
        return self._ax(self.b[1], buf, 1, base_backward)

    def to_ortho(self, vhat):
        buf = self._ax(self.b[0], vhat, 0, base_to_ortho)
        return self._ax(self.b[1], buf, 1, base_to_ortho)

    def from_ortho(self, p):
        buf = self._ax(self.b[0], p, 0, base_from_ortho)
        return self._ax(self.b[1], buf, 1, base_from_ortho)

    def gradient(self, vhat, deriv, scale=None):
        buf = _moveaxis_apply(vhat, 0, lambda x: base_differentiate(self.b[0], x, deriv[0]))
        out = _moveaxis_apply(buf, 1, lambda x: base_differentiate(self.b[1], x, deriv[1]))
        if scale is not None:
            sc = scale[0] ** deriv[0] * scale[1] ** deriv[1]
            if np.iscomplexobj(out):
                # num-complex 0.4 Complex/Complex division with im = 0
                re = (out.real * sc + out.imag * 0.0) / (sc * sc + 0.0 * 0.0)
                im = (out.imag * sc - out.real * 0.0) / (sc * sc + 0.0 * 0.0)
                out = re + 1j * im
            else:
                out = out / sc
        return out


class Field2:
    """src/field.rs:67-163"""

    def __init__(self, space):
        self.space = space
        self.v = np.zeros(space.shape_phys())
        dt = complex if space.is_complex else float
        self.vhat = np.zeros(space.shape_spec(), dtype=dt)
        self.x = [space.b[0].coords(), space.b[1].coords()]
        self.dx = [self._dx(self.x[0], space.b[0].is_fourier), self._dx(self.x[1], space.b[1].is_fourier)]

    @staticmethod
    def _dx(x, periodic):  # field.rs:135-163
        if periodic:
            return np.full(len(x), x[2] - x[1])
        n = len(x)
        dx = np.zeros(n)
        for i in range(n):
            left = x[0] if i == 0 else (x[i] + x[i - 1]) / 2.0
            right = x[n - 1] if i == n - 1 else (x[i + 1] + x[i]) / 2.0
            dx[i] = right - left
        return dx

    def forward(self):
        self.vhat = self.space.forward(self.v)

    def backward(self):
        self.v = self.space.backward(self.vhat)

    def to_ortho(self):
        return self.space.to_ortho(self.vhat)

    def from_ortho(self, p):
        self.vhat = self.space.from_ortho(p)

    def gradient(self, deriv, scale=None):
        return self.space.gradient(self.vhat, deriv, scale)

    # src/field/average.rs:25-57
    def average_axis(self, axis):
        length = abs(self.x[axis][-1] - self.x[axis][0])
        w = self.dx[axis] / length
        if axis == 0:
            return (self.v * w[:, None]).sum(axis=0)
        return (self.v * w[None, :]).sum(axis=1)

    def average(self):
        length = abs(self.x[1][-1] - self.x[1][0])
        avg_x = self.average_axis(0) * self.dx[1] / length
        return avg_x.sum()

    # ingredients field.rs:194-252
    def ingredients_for_hholtz(self, axis):
        b = self.space.b[axis]
        mass = b.mass()
        peye = b.laplace_inv_eye()
        pinv = peye @ b.laplace_inv()
        if b.kind == "cheb":
            ms = mass[:, 2:]
            mat_a, mat_b = pinv @ ms, peye @ ms
        elif b.kind == "fo":
            mat_a, mat_b = mass, b.laplace()
        else:
            mat_a, mat_b = pinv @ mass, peye @ mass
        precond = None if b.kind == "fo" else pinv
        return mat_a, mat_b, precond

    def ingredients_for_poisson(self, axis):
        a, b, p = self.ingredients_for_hholtz(axis)
        return a, b, p, self.space.b[axis].kind == "fo"


# ----------------------------------------------------------------------------------
# Solvers  (src/solver/*.rs)
# ----------------------------------------------------------------------------------


def diag_of(a, off):  # utils.rs:17-43
    return np.array(np.diagonal(a, off), copy=True)


class Fdma:
    """fdma.rs:11-118; diagonals at offsets -2,0,2,4."""

    def __init__(self, a=None, diags=None):
        if diags is not None:
            self.low, self.dia, self.up1, self.up2 = [np.array(d, dtype=float, copy=True) for d in diags]
        else:
            self.low, self.dia = diag_of(a, -2), diag_of(a, 0)
            self.up1, self.up2 = diag_of(a, 2), diag_of(a, 4)
        self.n = len(self.dia)

    def add_scaled(self, other, l):  # fdma.rs:187-227 : self + other*l
        return Fdma(diags=(self.low + other.low * l, self.dia + other.dia * l,
                           self.up1 + other.up1 * l, self.up2 + other.up2 * l))

    def sweep(self):  # fdma.rs:73-82
        n = self.n
        for i in range(2, n):
            self.low[i - 2] /= self.dia[i - 2]
            self.dia[i] -= self.low[i - 2] * self.up1[i - 2]
            if i < n - 2:
                self.up1[i] -= self.low[i - 2] * self.up2[i - 2]
        return self

    def solve(self, x):  # fdma.rs:101-118, last axis, returns new array
        x = np.array(x, copy=True)
        n = self.n
        for i in range(2, n):
            x[..., i] = x[..., i] - x[..., i - 2] * self.low[i - 2]
        x[..., n - 1] = x[..., n - 1] / self.dia[n - 1]
        x[..., n - 2] = x[..., n - 2] / self.dia[n - 2]
        x[..., n - 3] = (x[..., n - 3] - x[..., n - 1] * self.up1[n - 3]) / self.dia[n - 3]
        x[..., n - 4] = (x[..., n - 4] - x[..., n - 2] * self.up1[n - 4]) / self.dia[n - 4]
        for i in range(n - 5, -1, -1):
            x[..., i] = (x[..., i] - x[..., i + 2] * self.up1[i] - x[..., i + 4] * self.up2[i]) / self.dia[i]
        return x


class MatVecFdma:
    """matvec.rs:142-193 (offsets -2,0,2,4 of an m x n matrix)."""

    def __init__(self, a):
        m = a.shape[0]
        self.m = m
        self.low, self.dia = np.zeros(m), np.zeros(m)
        self.up1, self.up2 = np.zeros(m), np.zeros(m)
        for i in range(m):
            self.dia[i] = a[i, i]
            if i > 1:
                self.low[i] = a[i, i - 2]
            if i < m - 2:
                self.up1[i] = a[i, i + 2]
            if i < m - 4:
                self.up2[i] = a[i, i + 4]

    def apply(self, x):  # last axis
        m = self.m
        out = np.zeros(x.shape[:-1] + (m,), dtype=x.dtype)
        for i in range(m):
            o = x[..., i] * self.dia[i]
            if i > 1:
                o = o + x[..., i - 2] * self.low[i]
            if i < m - 2:
                o = o + x[..., i + 2] * self.up1[i]
            if i < m - 4:
                o = o + x[..., i + 4] * self.up2[i]
            out[..., i] = o
        return out


def eig_sorted(a):
    """utils.rs:66-98: LAPACK dgeev, real parts, sorted descending, Q^-1 = inv(Q)."""
    w, v = np.linalg.eig(a)
    w, v = w.real.copy(), v.real.copy()
    perm = np.argsort(w, kind="stable")[::-1]
    w, v = w[perm], v[:, perm]
    return w, v, np.linalg.inv(v)


class FdmaTensor2:
    """fdma_tensor.rs:105-153, 185-234 for N = 2."""

    def __init__(self, a, c, is_diag, alpha):
        if is_diag[0]:
            self.lam = diag_of(a[0], 0)
            self.fwd = self.bwd = None
        else:
            cinv = np.linalg.inv(c[0])
            lam, q, p = eig_sorted(cinv @ a[0])
            self.lam, self.fwd, self.bwd = lam, p @ cinv, q
        self.fa, self.fc = Fdma(a[1]), Fdma(c[1])
        self.alpha = alpha

    def solve(self, rhs):
        out = self.fwd @ rhs if self.fwd is not None else np.array(rhs, copy=True)
        for i in range(out.shape[0]):
            l = self.lam[i] + self.alpha
            f = self.fa.add_scaled(self.fc, l).sweep()
            out[i] = f.solve(out[i])
        if self.bwd is not None:
            out = self.bwd @ out
        return out


class HholtzAdi2:
    """hholtz_adi.rs:44-63, 106-131"""

    def __init__(self, field, c):
        self.solver, self.matvec = [], []
        for axis, ci in enumerate(c):
            a, b, p = field.ingredients_for_hholtz(axis)
            self.solver.append(Fdma(a - b * ci).sweep())
            self.matvec.append(None if p is None else MatVecFdma(p))

    def solve(self, rhs):
        r = rhs if self.matvec[0] is None else _moveaxis_apply(rhs, 0, self.matvec[0].apply)
        if self.matvec[1] is not None:
            r = self.matvec[1].apply(r)
        out = _moveaxis_apply(r, 0, self.solver[0].solve)
        return self.solver[1].solve(out)


class Hholtz2:
    """hholtz.rs:42-76, 159-197"""

    def __init__(self, field, c):
        lap, mass, isd, mv = [], [], [], []
        for axis, ci in enumerate(c):
            a, b, p, d = field.ingredients_for_poisson(axis)
            mass.append(a)
            lap.append(-1.0 * b * ci)
            mv.append(None if p is None else MatVecFdma(p))
            isd.append(d)
        self.matvec = mv
        self.t = FdmaTensor2(lap, mass, isd, 1.0)

    def solve(self, rhs):
        r = rhs if self.matvec[0] is None else _moveaxis_apply(rhs, 0, self.matvec[0].apply)
        if self.matvec[1] is not None:
            r = self.matvec[1].apply(r)
        return self.t.solve(r)


class Poisson2:
    """poisson.rs:50-90, 122-150"""

    def __init__(self, field, c):
        lap, mass, isd, mv = [], [], [], []
        for axis, ci in enumerate(c):
            a, b, p, d = field.ingredients_for_poisson(axis)
            mass.append(a)
            lap.append(b * ci)
            mv.append(None if p is None else MatVecFdma(p))
            isd.append(d)
        self.matvec = mv
        self.t = FdmaTensor2(lap, mass, isd, 0.0)
        if abs(self.t.lam[0]) < 1e-10:  # poisson.rs:80-83
            self.t.lam = self.t.lam - 1e-10

    def solve(self, rhs):
        r = rhs if self.matvec[0] is None else _moveaxis_apply(rhs, 0, self.matvec[0].apply)
        if self.matvec[1] is not None:
            r = self.matvec[1].apply(r)
        return self.t.solve(r)


# ----------------------------------------------------------------------------------
# Navier2D  (src/navier/navier.rs)
# ----------------------------------------------------------------------------------


def get_nu(ra, pr, height):  # navier.rs:46-49
    return np.sqrt(pr / (ra / height ** 3.0))


def get_ka(ra, pr, height):  # navier.rs:52-55
    return np.sqrt(1.0 / ((ra / height ** 3.0) * pr))


class Navier2D:
    def __init__(self, nx, ny, ra, pr, dt, aspect, adiabatic=True, periodic=False):
        self.periodic = periodic
        self.scale = [aspect, 1.0]
        self.nu = get_nu(ra, pr, self.scale[1] * 2.0)
        self.ka = get_ka(ra, pr, self.scale[1] * 2.0)
        self.ra, self.pr, self.dt, self.time = ra, pr, dt, 0.0
        B = Base
        bx = (lambda k: B("fo", nx)) if periodic else (lambda k: B(k, nx))
        self.ux = Field2(Space2(bx("cd"), B("cd", ny)))
        self.uy = Field2(Space2(bx("cd"), B("cd", ny)))
        tx = "cd" if (periodic or not adiabatic) else "cn"
        self.temp = Field2(Space2(bx(tx), B("cd", ny)))
        self.pres = [Field2(Space2(bx("cheb"), B("cheb", ny))), Field2(Space2(bx("cn"), B("cn", ny)))]
        self.field = Field2(Space2(bx("cheb"), B("cheb", ny)))
        s = self.scale
        cv = [dt * self.nu / s[0] ** 2.0, dt * self.nu / s[1] ** 2.0]
        ck = [dt * self.ka / s[0] ** 2.0, dt * self.ka / s[1] ** 2.0]
        if periodic:  # navier.rs:410-425
            self.solver = [Hholtz2(self.ux, cv), Hholtz2(self.uy, cv), Hholtz2(self.temp, ck)]
        else:  # navier.rs:250-265
            self.solver = [HholtzAdi2(self.ux, cv), HholtzAdi2(self.uy, cv), HholtzAdi2(self.temp, ck)]
        self.solver.append(Poisson2(self.pres[1], [1.0 / s[0] ** 2.0, 1.0 / s[1] ** 2.0]))
        for f in (self.temp, self.ux, self.uy, self.pres[0]):  # _scale navier.rs:500-515
            f.x = [f.x[0] * s[0], f.x[1] * s[1]]
            f.dx = [f.dx[0] * s[0], f.dx[1] * s[1]]
        self.fieldbc = self._bc_rbc(nx, ny)
        self.dealias = True

    def _bc_rbc(self, nx, ny):  # navier.rs:314-332, 474-492
        if self.periodic:
            xb = Base("fo", nx)
            fb = Field2(Space2(xb, Base("cdbc", ny)))
            bc = np.zeros((nx, 2))
            bc[:, 0], bc[:, 1] = 0.5, -0.5
            fb.vhat = np.fft.rfft(bc, axis=0)
        else:
            xb = Base("cheb", nx)
            fb = Field2(Space2(xb, Base("cdbc", ny)))
            bc = np.zeros((nx, 2))
            bc[:, 0], bc[:, 1] = 0.5, -0.5
            fb.vhat = cheb_forward(bc.T).T
        fb.backward()
        fb.forward()
        return fb

    # set_velocity navier.rs:927-930 ; apply_sin_cos/apply_cos_sin :1035-1076
    def set_velocity(self, amp, m, n):
        for f, a, sc in ((self.ux, amp, True), (self.uy, -amp, False)):
            x, y = f.x
            x = (x - x[0]) / (x[-1] - x[0])
            y = (y - y[0]) / (y[-1] - y[0])
            ax, ay = np.pi * m, np.pi * n
            if sc:
                f.v = a * np.sin(ax * x)[:, None] * np.cos(ay * y)[None, :]
            else:
                f.v = a * np.cos(ax * x)[:, None] * np.sin(ay * y)[None, :]
            f.forward()

    def set_temperature(self, amp, m, n):  # navier.rs:935-937
        f = self.temp
        x, y = f.x
        x = (x - x[0]) / (x[-1] - x[0])
        y = (y - y[0]) / (y[-1] - y[0])
        f.v = -amp * np.cos(np.pi * m * x)[:, None] * np.sin(np.pi * n * y)[None, :]
        f.forward()

    def conv_term(self, fld, u, deriv):  # conv_term.rs:22-42
        self.field.vhat = fld.gradient(deriv, self.scale)
        self.field.backward()
        return u * self.field.v

    def _finish_conv(self, conv):  # navier.rs:562-568
        self.field.v = conv
        self.field.forward()
        if self.dealias:  # navier.rs:1022-1032
            s0, s1 = self.field.vhat.shape
            self.field.vhat[s0 * 2 // 3:, :] = 0
            self.field.vhat[:, s1 * 2 // 3:] = 0
        return np.array(self.field.vhat, copy=True)

    def conv_temp(self, ux, uy):  # navier.rs:538-568
        conv = self.conv_term(self.temp, ux, [1, 0])
        conv = conv + self.conv_term(self.temp, uy, [0, 1])
        conv = conv + self.conv_term(self.fieldbc, ux, [1, 0])
        conv = conv + self.conv_term(self.fieldbc, uy, [0, 1])
        return self._finish_conv(conv)

    def conv_u(self, f, ux, uy):  # navier.rs:571-616
        conv = self.conv_term(f, ux, [1, 0])
        conv = conv + self.conv_term(f, uy, [0, 1])
        return self._finish_conv(conv)

    def solve_ux(self, ux, uy):  # navier.rs:622-633
        rhs = self.ux.to_ortho()
        rhs = rhs - self.pres[0].gradient([1, 0], self.scale) * self.dt
        rhs = rhs - self.conv_u(self.ux, ux, uy) * self.dt
        self.ux.vhat = self.solver[0].solve(rhs)

    def solve_uy(self, ux, uy, buoy):  # navier.rs:636-654
        rhs = self.uy.to_ortho()
        rhs = rhs - self.pres[0].gradient([0, 1], self.scale) * self.dt
        rhs = rhs + buoy * self.dt
        rhs = rhs - self.conv_u(self.uy, ux, uy) * self.dt
        self.uy.vhat = self.solver[1].solve(rhs)

    def solve_temp(self, ux, uy):  # navier.rs:660-674
        rhs = self.temp.to_ortho()
        rhs = rhs + self.fieldbc.gradient([2, 0], self.scale) * self.dt * self.ka
        rhs = rhs + self.fieldbc.gradient([0, 2], self.scale) * self.dt * self.ka
        rhs = rhs - self.conv_temp(ux, uy) * self.dt
        self.temp.vhat = self.solver[2].solve(rhs)

    def divergence(self):  # navier.rs:698-703
        return self.ux.gradient([1, 0], self.scale) + self.uy.gradient([0, 1], self.scale)

    def solve_pres(self, div):  # navier.rs:710-715
        self.pres[1].vhat = self.solver[3].solve(div)
        self.pres[1].vhat[0, 0] = 0

    def project_velocity(self, c):  # navier.rs:683-695
        dpdx = self.pres[1].gradient([1, 0], self.scale)
        dpdy = self.pres[1].gradient([0, 1], self.scale)
        ux_old, uy_old = self.ux.vhat.copy(), self.uy.vhat.copy()
        self.ux.from_ortho(dpdx)
        self.uy.from_ortho(dpdy)
        self.ux.vhat = self.ux.vhat * (-c) + ux_old
        self.uy.vhat = self.uy.vhat * (-c) + uy_old

    def update_pres(self, div):  # navier.rs:717-721
        self.pres[0].vhat = self.pres[0].vhat - div * self.nu
        self.pres[0].vhat = self.pres[0].vhat + self.pres[1].to_ortho() * (1.0 / self.dt)

    def update(self):  # navier.rs:737-765
        that = self.temp.to_ortho() + self.fieldbc.to_ortho()
        self.ux.backward()
        self.uy.backward()
        ux, uy = self.ux.v.copy(), self.uy.v.copy()
        self.solve_ux(ux, uy)
        self.solve_uy(ux, uy, that)
        div = self.divergence()
        self.solve_pres(div)
        self.project_velocity(1.0)
        self.update_pres(div)
        self.solve_temp(ux, uy)
        self.time += self.dt

    # diagnostics src/navier/functions.rs:12-101, navier.rs:869-879
    def eval_nu(self):
        self.field.vhat = self.temp.to_ortho() + self.fieldbc.to_ortho()
        dtdz = self.field.gradient([0, 1], None) * -1.0
        dtdz = dtdz * (1.0 / (self.scale[1] / 2.0))
        self.field.vhat = dtdz
        self.field.backward()
        xa = self.field.average_axis(0)
        return (xa[-1] + xa[0]) / 2.0

    def eval_nuvol(self):
        self.field.vhat = self.temp.to_ortho() + self.fieldbc.to_ortho()
        self.field.backward()
        self.uy.backward()
        uy_temp = self.field.v * self.uy.v
        dtdz = self.field.gradient([0, 1], None) / (self.scale[1] * -1.0)
        self.field.vhat = dtdz
        self.field.backward()
        self.field.v = (self.field.v + uy_temp / self.ka) * 2.0 * self.scale[1]
        return self.field.average()

    def eval_re(self):
        self.ux.backward()
        self.uy.backward()
        ekin = self.ux.v ** 2 + self.uy.v ** 2
        self.field.v = np.sqrt(ekin) * (2.0 * self.scale[1] / self.nu)
        return self.field.average()

    def div_norm(self):
        d = self.divergence()
        return float(np.sqrt(np.sum(np.abs(d) ** 2)))
