"""Scattering of a plane wave by a dielectric circular cylinder (the Mie series
of the 2-D problem; Bohren & Huffman, "Absorption and Scattering of Light by
Small Particles", sect. 8.4).  Used to pin the far fields physically.

With incidence along +x, relative index m and size parameter ka, the
scattered far field is proportional to T(phi) = sum_n c_n e^{i n phi}, phi
measured from the forward direction, with
  E_z (TM):  c_n = [J_n(mka) J_n'(ka) - m J_n'(mka) J_n(ka)] / [J_n(mka) H_n'(ka) - m J_n'(mka) H_n(ka)]
  H_z (TE):  c_n = [m J_n(mka) J_n'(ka) - J_n'(mka) J_n(ka)] / [m J_n(mka) H_n'(ka) - J_n'(mka) H_n(ka)]
(H = H^(1), the outgoing Hankel function for e^{-i w t}).
"""
import numpy as np
from scipy.special import h1vp, hankel1, jv, jvp


def pattern(ka, m, pol, deg=np.arange(360)):
    """|T(phi)|^2 normalised to its maximum, phi in degrees from the forward direction."""
    phi = np.deg2rad(deg)
    nmax = int(ka + 4 * ka ** (1 / 3) + 10)
    T = np.zeros(len(phi), complex)
    mka = m * ka
    for n in range(-nmax, nmax + 1):
        if pol == "tm":
            num = jv(n, mka) * jvp(n, ka) - m * jvp(n, mka) * jv(n, ka)
            den = jv(n, mka) * h1vp(n, ka) - m * jvp(n, mka) * hankel1(n, ka)
        else:
            num = m * jv(n, mka) * jvp(n, ka) - jvp(n, mka) * jv(n, ka)
            den = m * jv(n, mka) * h1vp(n, ka) - jvp(n, mka) * hankel1(n, ka)
        T += (num / den) * np.exp(1j * n * phi)
    p = np.abs(T) ** 2
    return p / p.max()


def compare(ff, ka, m, pol):
    """(error vs the series, error vs the other polarisation's series, mirror asymmetry)
    of a far field in the reference's convention: its NTFF integrals use the phase
    e^{+i k r.r'} (ntffTM.c:88-110) with fields that vary as e^{-i w t}, so index
    `ang` holds the direction ang + 180 degrees, and the forward direction of +x
    incidence sits at index 180."""
    p = np.abs(np.roll(ff, 180)) ** 2
    p = p / p.max()
    q = pattern(ka, m, pol)
    r = pattern(ka, m, "te" if pol == "tm" else "tm")
    err = np.linalg.norm(p - q) / np.linalg.norm(q)
    other = np.linalg.norm(p - r) / np.linalg.norm(r)
    asym = np.linalg.norm(p - np.roll(p[::-1], 1)) / np.linalg.norm(p)
    return err, other, asym
