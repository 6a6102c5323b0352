"""Gate matrices of ket (src/gates.ts:1-208), built on the host.

A Gate2x2 is a 4-tuple (a, b, c, d) of Python complex for [[a, b], [c, d]];
a Gate4x4 is a 16-tuple in row-major order over |00>,|01>,|10>,|11> with the
first qubit argument as the MSB (src/statevector.ts:30-38).  Entries are
evaluated with the same operations as gates.ts, so they are the same doubles.
"""
import math

sq2 = 1 / math.sqrt(2)
ONE = complex(1, 0)
ZERO = complex(0, 0)


def c(re, im=0.0):
    return complex(re, im)


Id = (ONE, ZERO, ZERO, ONE)
H = (c(sq2), c(sq2), c(sq2), c(-sq2))
X = (ZERO, ONE, ONE, ZERO)
Y = (ZERO, c(0, -1), c(0, 1), ZERO)
Z = (ONE, ZERO, ZERO, c(-1))
S = (ONE, ZERO, ZERO, c(0, 1))
Si = (ONE, ZERO, ZERO, c(0, -1))
T = (ONE, ZERO, ZERO, c(sq2, sq2))
Ti = (ONE, ZERO, ZERO, c(sq2, -sq2))
V = (c(.5, .5), c(.5, -.5), c(.5, -.5), c(.5, .5))
Vi = (c(.5, -.5), c(.5, .5), c(.5, .5), c(.5, -.5))


def Rx(theta):
    co, si = math.cos(theta / 2), math.sin(theta / 2)
    return (c(co), c(0, -si), c(0, -si), c(co))


def Ry(theta):
    co, si = math.cos(theta / 2), math.sin(theta / 2)
    return (c(co, 0), c(-si), c(si), c(co))


def Rz(theta):
    co, si = math.cos(theta / 2), math.sin(theta / 2)
    return (c(co, -si), ZERO, ZERO, c(co, si))


R2 = Rz(math.pi / 2)
R4 = Rz(math.pi / 4)
R8 = Rz(math.pi / 8)


def U3(theta, phi, lam):
    co, si = math.cos(theta / 2), math.sin(theta / 2)
    return (c(co), c(-si * math.cos(lam), -si * math.sin(lam)),
            c(si * math.cos(phi), si * math.sin(phi)), c(co * math.cos(phi + lam), co * math.sin(phi + lam)))


def U2(phi, lam):
    return U3(math.pi / 2, phi, lam)


def U1(lam):
    return (ONE, ZERO, ZERO, c(math.cos(lam), math.sin(lam)))


def Xx(theta):
    co, ni = c(math.cos(theta / 2)), c(0, -math.sin(theta / 2))
    return (co, ZERO, ZERO, ni,
            ZERO, co, ni, ZERO,
            ZERO, ni, co, ZERO,
            ni, ZERO, ZERO, co)


def Yy(theta):
    co = c(math.cos(theta / 2))
    ni = c(0, -math.sin(theta / 2))
    pi = c(0, math.sin(theta / 2))
    return (co, ZERO, ZERO, pi,
            ZERO, co, ni, ZERO,
            ZERO, ni, co, ZERO,
            pi, ZERO, ZERO, co)


def Zz(theta):
    co, si = math.cos(theta / 2), math.sin(theta / 2)
    m, p = c(co, -si), c(co, si)
    return (m, ZERO, ZERO, ZERO,
            ZERO, p, ZERO, ZERO,
            ZERO, ZERO, p, ZERO,
            ZERO, ZERO, ZERO, m)


def Xy(theta):
    co, is_ = c(math.cos(theta / 2)), c(0, math.sin(theta / 2))
    return (ONE, ZERO, ZERO, ZERO,
            ZERO, co, is_, ZERO,
            ZERO, is_, co, ZERO,
            ZERO, ZERO, ZERO, ONE)


ISwap = Xy(math.pi)
SrSwap = Xy(math.pi / 2)


def Gpi(phi):
    co, si = math.cos(phi), math.sin(phi)
    return (ZERO, c(co, -si), c(co, si), ZERO)


def Gpi2(phi):
    co, si = math.cos(phi), math.sin(phi)
    return (c(sq2), c(-sq2 * si, -sq2 * co), c(sq2 * si, -sq2 * co), c(sq2))


def Ms(phi0, phi1):
    sp, cp = math.sin(phi0 + phi1), math.cos(phi0 + phi1)
    sd, cd = math.sin(phi0 - phi1), math.cos(phi0 - phi1)
    return (c(sq2), ZERO, ZERO, c(-sp * sq2, -cp * sq2),
            ZERO, c(sq2), c(-sd * sq2, -cd * sq2), ZERO,
            ZERO, c(sd * sq2, -cd * sq2), c(sq2), ZERO,
            c(sp * sq2, -cp * sq2), ZERO, ZERO, c(sq2))


# ---------------------------------------------------------------------------
# Derivatives of the parametrised matrices with respect to each angle, used
# by the adjoint gradient (d/dtheta of the constructors above).
# ---------------------------------------------------------------------------
def dRx(theta):
    co, si = math.cos(theta / 2), math.sin(theta / 2)
    return (c(-si / 2), c(0, -co / 2), c(0, -co / 2), c(-si / 2))


def dRy(theta):
    co, si = math.cos(theta / 2), math.sin(theta / 2)
    return (c(-si / 2), c(-co / 2), c(co / 2), c(-si / 2))


def dRz(theta):
    co, si = math.cos(theta / 2), math.sin(theta / 2)
    return (c(-si / 2, -co / 2), ZERO, ZERO, c(-si / 2, co / 2))


def dU1(lam):
    return (ZERO, ZERO, ZERO, c(-math.sin(lam), math.cos(lam)))


def dU3(theta, phi, lam, which):
    co, si = math.cos(theta / 2), math.sin(theta / 2)
    if which == 0:   # d/dtheta
        return (c(-si / 2), c(-co / 2 * math.cos(lam), -co / 2 * math.sin(lam)),
                c(co / 2 * math.cos(phi), co / 2 * math.sin(phi)),
                c(-si / 2 * math.cos(phi + lam), -si / 2 * math.sin(phi + lam)))
    if which == 1:   # d/dphi
        return (ZERO, ZERO, c(-si * math.sin(phi), si * math.cos(phi)),
                c(-co * math.sin(phi + lam), co * math.cos(phi + lam)))
    return (ZERO, c(si * math.sin(lam), -si * math.cos(lam)), ZERO,   # d/dlambda
            c(-co * math.sin(phi + lam), co * math.cos(phi + lam)))


def dU2(phi, lam, which):
    return dU3(math.pi / 2, phi, lam, which + 1)


def dGpi(phi):
    co, si = math.cos(phi), math.sin(phi)
    return (ZERO, c(-si, -co), c(-si, co), ZERO)


def dGpi2(phi):
    co, si = math.cos(phi), math.sin(phi)
    return (ZERO, c(-sq2 * co, sq2 * si), c(sq2 * co, sq2 * si), ZERO)


def dXx(theta):
    co, ni = c(-math.sin(theta / 2) / 2), c(0, -math.cos(theta / 2) / 2)
    return (co, ZERO, ZERO, ni, ZERO, co, ni, ZERO, ZERO, ni, co, ZERO, ni, ZERO, ZERO, co)


def dYy(theta):
    co = c(-math.sin(theta / 2) / 2)
    ni = c(0, -math.cos(theta / 2) / 2)
    pi = c(0, math.cos(theta / 2) / 2)
    return (co, ZERO, ZERO, pi, ZERO, co, ni, ZERO, ZERO, ni, co, ZERO, pi, ZERO, ZERO, co)


def dZz(theta):
    co, si = math.cos(theta / 2), math.sin(theta / 2)
    m, p = c(-si / 2, -co / 2), c(-si / 2, co / 2)
    return (m, ZERO, ZERO, ZERO, ZERO, p, ZERO, ZERO, ZERO, ZERO, p, ZERO, ZERO, ZERO, ZERO, m)


def dXy(theta):
    co, is_ = c(-math.sin(theta / 2) / 2), c(0, math.cos(theta / 2) / 2)
    return (ZERO, ZERO, ZERO, ZERO, ZERO, co, is_, ZERO, ZERO, is_, co, ZERO, ZERO, ZERO, ZERO, ZERO)


def dMs(phi0, phi1, which):
    sp, cp = math.sin(phi0 + phi1), math.cos(phi0 + phi1)
    sd, cd = math.sin(phi0 - phi1), math.cos(phi0 - phi1)
    s = 1.0 if which == 0 else -1.0          # d(phi0 - phi1)/d(which)
    return (ZERO, ZERO, ZERO, c(-cp * sq2, sp * sq2),
            ZERO, ZERO, c(-s * cd * sq2, s * sd * sq2), ZERO,
            ZERO, c(s * cd * sq2, s * sd * sq2), ZERO, ZERO,
            c(cp * sq2, sp * sq2), ZERO, ZERO, ZERO)
