"""The parametrised gate matrices of gates.py, and their angle derivatives,
evaluated for many angle values at once (numpy arrays of shape (B,)), with
the same expressions in the same order as the scalar constructors.  Used to
fill the matrix arrays of many parameter sets of one ansatz (the batched
energy / gradient path).  Each function returns a complex array of shape
(B, 4) (2x2, row-major) or (B, 16) (4x4)."""
import numpy as np

sq2 = 1 / np.sqrt(2)


def _c(re, im=None):
    re = np.asarray(re, dtype=np.float64)
    return re + 0j if im is None else re + 1j * np.asarray(im, dtype=np.float64)


def _stack(*cols):
    B = max(np.shape(c)[0] if np.ndim(c) else 1 for c in cols)
    return np.stack([np.broadcast_to(np.asarray(c, dtype=np.complex128), (B,)) for c in cols], axis=1)


def _z(x):
    """real zeros shaped like x (the +0.0 of a complex literal)"""
    return np.zeros(np.shape(x))


def _zc(x):
    return np.zeros(np.shape(x), dtype=np.complex128)


def _oc(x):
    return np.ones(np.shape(x), dtype=np.complex128)


def Rx(t):
    co, si = np.cos(t / 2), np.sin(t / 2)
    return _stack(_c(co), _c(_z(si), -si), _c(_z(si), -si), _c(co))


def Ry(t):
    co, si = np.cos(t / 2), np.sin(t / 2)
    return _stack(_c(co, _z(co)), _c(-si), _c(si), _c(co))


def Rz(t):
    co, si = np.cos(t / 2), np.sin(t / 2)
    return _stack(_c(co, -si), _zc(co), _zc(co), _c(co, si))


def U3(th, ph, la):
    co, si = np.cos(th / 2), np.sin(th / 2)
    return _stack(_c(co), _c(-si * np.cos(la), -si * np.sin(la)), _c(si * np.cos(ph), si * np.sin(ph)),
                  _c(co * np.cos(ph + la), co * np.sin(ph + la)))


def U2(ph, la):
    return U3(np.full(np.shape(ph), np.pi / 2), ph, la)


def U1(la):
    return _stack(_oc(la), _zc(la), _zc(la), _c(np.cos(la), np.sin(la)))


def Gpi(ph):
    co, si = np.cos(ph), np.sin(ph)
    return _stack(_zc(co), _c(co, -si), _c(co, si), _zc(co))


def Gpi2(ph):
    co, si = np.cos(ph), np.sin(ph)
    return _stack(_c(sq2 + _z(co)), _c(-sq2 * si, -sq2 * co), _c(sq2 * si, -sq2 * co), _c(sq2 + _z(co)))


def Xx(t):
    co, ni = _c(np.cos(t / 2)), _c(_z(t), -np.sin(t / 2))
    z = _zc(co)
    return _stack(co, z, z, ni, z, co, ni, z, z, ni, co, z, ni, z, z, co)


def Yy(t):
    co = _c(np.cos(t / 2))
    ni = _c(_z(t), -np.sin(t / 2))
    pi = _c(_z(t), np.sin(t / 2))
    z = _zc(co)
    return _stack(co, z, z, pi, z, co, ni, z, z, ni, co, z, pi, z, z, co)


def Zz(t):
    co, si = np.cos(t / 2), np.sin(t / 2)
    m, p = _c(co, -si), _c(co, si)
    z = _zc(m)
    return _stack(m, z, z, z, z, p, z, z, z, z, p, z, z, z, z, m)


def Xy(t):
    co, is_ = _c(np.cos(t / 2)), _c(_z(t), np.sin(t / 2))
    z, o = _zc(co), _oc(co)
    return _stack(o, z, z, z, z, co, is_, z, z, is_, co, z, z, z, z, o)


def Ms(p0, p1):
    sp, cp = np.sin(p0 + p1), np.cos(p0 + p1)
    sd, cd = np.sin(p0 - p1), np.cos(p0 - p1)
    s2 = _c(sq2 + _z(sp))
    z = _zc(s2)
    return _stack(s2, z, z, _c(-sp * sq2, -cp * sq2),
                  z, s2, _c(-sd * sq2, -cd * sq2), z,
                  z, _c(sd * sq2, -cd * sq2), s2, z,
                  _c(sp * sq2, -cp * sq2), z, z, s2)


def dRx(t):
    co, si = np.cos(t / 2), np.sin(t / 2)
    return _stack(_c(-si / 2), _c(_z(co), -co / 2), _c(_z(co), -co / 2), _c(-si / 2))


def dRy(t):
    co, si = np.cos(t / 2), np.sin(t / 2)
    return _stack(_c(-si / 2), _c(-co / 2), _c(co / 2), _c(-si / 2))


def dRz(t):
    co, si = np.cos(t / 2), np.sin(t / 2)
    return _stack(_c(-si / 2, -co / 2), _zc(co), _zc(co), _c(-si / 2, co / 2))


def dU1(la):
    z = _zc(la)
    return _stack(z, z, z, _c(-np.sin(la), np.cos(la)))


def dU3(th, ph, la, which):
    co, si = np.cos(th / 2), np.sin(th / 2)
    z = _zc(co)
    if which == 0:
        return _stack(_c(-si / 2), _c(-co / 2 * np.cos(la), -co / 2 * np.sin(la)),
                      _c(co / 2 * np.cos(ph), co / 2 * np.sin(ph)),
                      _c(-si / 2 * np.cos(ph + la), -si / 2 * np.sin(ph + la)))
    if which == 1:
        return _stack(z, z, _c(-si * np.sin(ph), si * np.cos(ph)), _c(-co * np.sin(ph + la), co * np.cos(ph + la)))
    return _stack(z, _c(si * np.sin(la), -si * np.cos(la)), z, _c(-co * np.sin(ph + la), co * np.cos(ph + la)))


def dU2(ph, la, which):
    return dU3(np.full(np.shape(ph), np.pi / 2), ph, la, which + 1)


def dGpi(ph):
    co, si = np.cos(ph), np.sin(ph)
    z = _zc(co)
    return _stack(z, _c(-si, -co), _c(-si, co), z)


def dGpi2(ph):
    co, si = np.cos(ph), np.sin(ph)
    z = _zc(co)
    return _stack(z, _c(-sq2 * co, sq2 * si), _c(sq2 * co, sq2 * si), z)


def dXx(t):
    co, ni = _c(-np.sin(t / 2) / 2), _c(_z(t), -np.cos(t / 2) / 2)
    z = _zc(co)
    return _stack(co, z, z, ni, z, co, ni, z, z, ni, co, z, ni, z, z, co)


def dYy(t):
    co = _c(-np.sin(t / 2) / 2)
    ni = _c(_z(t), -np.cos(t / 2) / 2)
    pi = _c(_z(t), np.cos(t / 2) / 2)
    z = _zc(co)
    return _stack(co, z, z, pi, z, co, ni, z, z, ni, co, z, pi, z, z, co)


def dZz(t):
    co, si = np.cos(t / 2), np.sin(t / 2)
    m, p = _c(-si / 2, -co / 2), _c(-si / 2, co / 2)
    z = _zc(m)
    return _stack(m, z, z, z, z, p, z, z, z, z, p, z, z, z, z, m)


def dXy(t):
    co, is_ = _c(-np.sin(t / 2) / 2), _c(_z(t), np.cos(t / 2) / 2)
    z = _zc(co)
    return _stack(z, z, z, z, z, co, is_, z, z, is_, co, z, z, z, z, z)


def dMs(p0, p1, which):
    sp, cp = np.sin(p0 + p1), np.cos(p0 + p1)
    sd, cd = np.sin(p0 - p1), np.cos(p0 - p1)
    s = 1.0 if which == 0 else -1.0
    z = _zc(sp)
    return _stack(z, z, z, _c(-cp * sq2, sp * sq2),
                  z, z, _c(-s * cd * sq2, s * sd * sq2), z,
                  z, _c(s * cd * sq2, s * sd * sq2), z, z,
                  _c(cp * sq2, sp * sq2), z, z, z)


# gate method name -> (matrix constructor, number of angles)
MATRIX = {"rx": (Rx, 1), "ry": (Ry, 1), "rz": (Rz, 1), "vz": (Rz, 1), "u1": (U1, 1), "p": (U1, 1), "u2": (U2, 2),
          "u3": (U3, 3), "gpi": (Gpi, 1), "gpi2": (Gpi2, 1), "xx": (Xx, 1), "yy": (Yy, 1), "zz": (Zz, 1),
          "xy": (Xy, 1), "ms": (Ms, 2), "crx": (Rx, 1), "cry": (Ry, 1), "crz": (Rz, 1), "cu1": (U1, 1),
          "cu2": (U2, 2), "cu3": (U3, 3)}


def derivative(name, vals, which):
    """d(matrix)/d(angle `which`) for gate method `name` at angles `vals` (arrays)."""
    if name in ("rx", "crx"):
        return dRx(vals[0])
    if name in ("ry", "cry"):
        return dRy(vals[0])
    if name in ("rz", "vz", "crz"):
        return dRz(vals[0])
    if name in ("u1", "p", "cu1"):
        return dU1(vals[0])
    if name in ("u2", "cu2"):
        return dU2(vals[0], vals[1], which)
    if name in ("u3", "cu3"):
        return dU3(vals[0], vals[1], vals[2], which)
    if name == "gpi":
        return dGpi(vals[0])
    if name == "gpi2":
        return dGpi2(vals[0])
    if name == "xx":
        return dXx(vals[0])
    if name == "yy":
        return dYy(vals[0])
    if name == "zz":
        return dZz(vals[0])
    if name == "xy":
        return dXy(vals[0])
    if name == "ms":
        return dMs(vals[0], vals[1], which)
    return None
