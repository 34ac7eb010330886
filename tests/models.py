"""Small model builders shared by the CPU oracle tests and the GPU parity tests.

The one- and two-orbital square lattices follow the reference's test builders
(tests/_tools.py:114-175); the exact Lindhard bubble follows
tests/_tools.py:245-283.  Everything is computed with the oracle (CPU)."""
import numpy as np

import oracle.pyrpa_oracle as O

RVEC_NN = np.array([[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0]], dtype=np.float64)


def one_orbital(Nx=4, Ny=4, Nz=1, Nw=16, temp=0.1, mu=0.0, tx=1.0, ty=0.5, U=0.0, J=0.0):
    ham_r = np.array([[[0.0]], [[-tx]], [[-tx]], [[-ty]], [[-ty]]], dtype=np.complex128)
    return _bundle(ham_r, RVEC_NN, 1, Nx, Ny, Nz, Nw, temp, mu, U, J)


def two_orbital(Nx=4, Ny=4, Nz=1, Nw=16, temp=0.1, mu=0.0, t1=1.0, t2=0.5, de=0.4, V=0.2,
                U=0.0, J=0.0):
    onsite = np.array([[0.0, V], [V, de]])
    hop = np.diag([-t1, -t2]).astype(np.complex128)
    ham_r = np.array([onsite, hop, hop, hop, hop], dtype=np.complex128)
    return _bundle(ham_r, RVEC_NN, 2, Nx, Ny, Nz, Nw, temp, mu, U, J)


def random_model(Norb, Nx, Ny, Nz, Nw, temp, mu, U, J, seed=0, amp=0.3, interlayer=0.2):
    """Seeded Hermitian hoppings to first neighbours (+ z hopping if Nz>1)."""
    rng = np.random.default_rng(seed)
    Rs = [(1, 0, 0), (0, 1, 0), (1, 1, 0), (1, -1, 0)]
    if Nz > 1:
        Rs += [(0, 0, 1)]
    rvec = [(0, 0, 0)]
    onsite = rng.uniform(-0.2, 0.2, size=(Norb, Norb)) + 1j * rng.uniform(-0.05, 0.05, size=(Norb, Norb))
    onsite = 0.5 * (onsite + onsite.conj().T)
    hams = [onsite]
    for R in Rs:
        h = rng.uniform(-amp, amp, size=(Norb, Norb)) + 1j * rng.uniform(-amp, amp, size=(Norb, Norb)) * 0.3
        if R[2] != 0:
            h = h * interlayer
        rvec += [R, tuple(-x for x in R)]
        hams += [h, h.conj().T]
    return _bundle(np.array(hams, dtype=np.complex128), np.array(rvec, dtype=np.float64), Norb,
                   Nx, Ny, Nz, Nw, temp, mu, U, J)


def _bundle(ham_r, rvec, Norb, Nx, Ny, Nz, Nw, temp, mu, U, J):
    klist, kmap, invk = O.gen_irr_k_TRS(Nx, Ny, Nz)
    hamk = O.gen_ham(klist, ham_r, rvec)
    eig, uni = O.get_eig(hamk)
    Gk = O.gen_Green0(eig, uni, mu, temp, Nw)
    olist, site = O.get_chi_orb_list(Norb, [Norb])
    Smat, Cmat = O.get_scmat(olist, site, U, J)
    return dict(Nx=Nx, Ny=Ny, Nz=Nz, Nw=Nw, temp=temp, mu=mu, ham_r=ham_r, rvec=rvec,
                klist=klist, kmap=kmap, invk=invk, hamk=hamk, eig=eig, uni=uni, Gk=Gk,
                olist=olist, site=site, Smat=Smat, Cmat=Cmat, Norb=Norb,
                plist=np.ones(Norb))


def fermi(e, temp, mu):
    x = (np.asarray(e) - mu) / temp
    return np.where(x > 0, np.exp(-x) / (1 + np.exp(-x)), 1 / (1 + np.exp(x)))


def lindhard_exact(st, mlist):
    """Exact Matsubara bubble of a 1-orbital model, chi = -T sum G G > 0
    (tests/_tools.py:245-283 with a=0)."""
    Nx, Ny = st['Nx'], st['Ny']
    ix, iy = np.meshgrid(np.arange(Nx), np.arange(Ny), indexing='ij')
    kf = np.stack([ix.ravel() / Nx, iy.ravel() / Ny, np.zeros(Nx * Ny)], axis=1)
    epsg = O.gen_ham(kf, st['ham_r'], st['rvec'])[:, 0, 0].real.reshape(Nx, Ny)
    temp, mu = st['temp'], st['mu']
    fg = fermi(epsg, temp, mu)
    out = np.zeros((len(st['klist']), len(mlist)), dtype=complex)
    for iq, q in enumerate(st['klist']):
        iqx = int(round(q[0] * Nx)) % Nx
        iqy = int(round(q[1] * Ny)) % Ny
        eq = np.roll(np.roll(epsg, -iqx, axis=0), -iqy, axis=1)
        fq = np.roll(np.roll(fg, -iqx, axis=0), -iqy, axis=1)
        for jm, m in enumerate(mlist):
            de = 2j * np.pi * m * temp + epsg - eq
            deg = (m == 0) & (np.abs(epsg - eq) < 1e-9)
            safe = np.where(np.abs(de) < 1e-30, 1.0, de)
            out[iq, jm] = np.where(deg, fg * (1 - fg) / temp, (fq - fg) / safe).mean()
    return out


def G0_at(st, Nw, mu=None):
    mu = st['mu'] if mu is None else mu
    return O.gen_Green0(st['eig'], st['uni'], mu, st['temp'], Nw)
