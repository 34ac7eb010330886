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


def random_model(Norb, Nx, Ny, Nz, Nw, temp, mu, U, J, seed=0, amp=0.3, interlayer=0.2, real=False):
    """Seeded Hermitian hoppings to first neighbours (+ z hopping if Nz>1).
    Complex by default (time reversal broken: the library and the oracle run
    the reference's algorithm on it all the same); real=True keeps the
    hoppings real, i.e. time-reversal symmetric as the reference's k/-k
    unfolding assumes physically."""
    rng = np.random.default_rng(seed)
    Rs = [(1, 0, 0), (0, 1, 0), (1, 1, 0), (1, -1, 0)]
    if Nz > 1:
        Rs += [(0, 0, 1)]
    rvec = [(0, 0, 0)]
    onsite = rng.uniform(-0.2, 0.2, size=(Norb, Norb)) + 1j * rng.uniform(-0.05, 0.05, size=(Norb, Norb))
    if real:
        onsite = onsite.real
    onsite = 0.5 * (onsite + onsite.conj().T)
    hams = [onsite]
    for R in Rs:
        h = rng.uniform(-amp, amp, size=(Norb, Norb)) + 1j * rng.uniform(-amp, amp, size=(Norb, Norb)) * 0.3
        if real:
            h = h.real
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


def bubble_exact(st, mlist):
    """Exact Matsubara bubble of a multi-orbital G0 from a band sum over the
    full mesh, in the layout of chi0 (Python [Nchi, Nchi, len(mlist), Nk]):
    chi_{(l1 l2)(m1 m2)}(q, i nu) = -(T/N) sum_k sum_n G_{l1 m1}(k+q, i w + i nu) G_{m2 l2}(k, i w)
      = (1/N) sum_k sum_ab U_{l1 a}(k+q) U*_{m1 a}(k+q) U_{m2 b}(k) U*_{l2 b}(k)
        [f(e_a(k+q)) - f(e_b(k))] / (i nu + e_b(k) - e_a(k+q)),
    -f'(e) = f (1 - f)/T at equal energies and nu = 0."""
    Nx, Ny, Nz = st['Nx'], st['Ny'], st['Nz']
    temp, mu = st['temp'], st['mu']
    ix, iy, iz = np.meshgrid(np.arange(Nx), np.arange(Ny), np.arange(Nz), indexing='ij')
    ix, iy, iz = ix.ravel(), iy.ravel(), iz.ravel()
    kf = np.stack([ix / Nx, iy / Ny, iz / Nz], axis=1)
    e, u = O.get_eig(O.gen_ham(kf, st['ham_r'], st['rvec']))
    U = np.transpose(u, (0, 2, 1))                      # U[k, orb, band]
    f = fermi(e, temp, mu)
    ol = np.asarray(st['olist']) - 1
    Nchi = len(ol)
    out = np.zeros((Nchi, Nchi, len(mlist), len(st['klist'])), dtype=complex)
    for iq, q in enumerate(st['klist']):
        qx, qy, qz = (int(round(q[0] * Nx)) % Nx, int(round(q[1] * Ny)) % Ny, int(round(q[2] * Nz)) % Nz)
        kq = ((ix + qx) % Nx) * Ny * Nz + ((iy + qy) % Ny) * Nz + (iz + qz) % Nz
        ea, fa, Ua = e[kq], f[kq], U[kq]
        A = Ua[:, :, None, :] * np.conj(Ua[:, None, :, :])     # [k, l1, m1, a]
        B = U[:, :, None, :] * np.conj(U[:, None, :, :])       # [k, m2, l2, b]
        for jm, m in enumerate(mlist):
            de = 2j * np.pi * m * temp + e[:, None, :] - ea[:, :, None]      # [k, a, b]
            deg = (m == 0) & (np.abs(e[:, None, :] - ea[:, :, None]) < 1e-9)
            safe = np.where(np.abs(de) < 1e-30, 1.0, de)
            r = np.where(deg, (f * (1 - f))[:, None, :] / temp, (fa[:, :, None] - f[:, None, :]) / safe)
            X = np.einsum('kxya,kzwb,kab->xyzw', A, B, r) / len(kf)   # [l1, m1, m2, l2]
            for l in range(Nchi):
                for mm in range(Nchi):
                    out[l, mm, jm, iq] = X[ol[l, 0], ol[mm, 0], ol[mm, 1], ol[l, 1]]
    return out


PAULI = {'x': np.array([[0, 1], [1, 0]], dtype=complex), 'y': np.array([[0, -1j], [1j, 0]]),
         'z': np.array([[1, 0], [0, -1]], dtype=complex)}


def spinful(st, lam=0.0, alpha=0.0, U=0.0, J=0.0):
    """The spinful version of a model on RVEC_NN (index s No + o, spin up
    first, as the reference driver's default slist / invs): H(k) for each
    spin, plus for two orbitals an atomic lam/2 L_z s_z term and a Rashba-like
    hopping i alpha (s_y along x, -s_x along y).  Every added term keeps time
    reversal, H(-k) = s_y H(k)^* s_y, the symmetry the spinful k / -k
    unfolding assumes.  Adds slist, invs and the SOC vertex Vmat(U, J)."""
    No = st['Norb']
    Ns = 2 * No
    hs = np.array([np.kron(np.eye(2), h) for h in st['ham_r']])
    if lam:
        assert No == 2
        Lz = np.array([[0, -1j], [1j, 0]])
        hs[0] += 0.5 * lam * np.kron(PAULI['z'], Lz)
    if alpha:
        for i, R in enumerate(st['rvec']):
            R = tuple(int(v) for v in R)
            if R == (1, 0, 0):
                hs[i] += np.kron(1j * alpha * PAULI['y'], np.eye(No))
            elif R == (-1, 0, 0):
                hs[i] += np.kron(1j * alpha * PAULI['y'], np.eye(No)).conj().T
            elif R == (0, 1, 0):
                hs[i] += np.kron(-1j * alpha * PAULI['x'], np.eye(No))
            elif R == (0, -1, 0):
                hs[i] += np.kron(-1j * alpha * PAULI['x'], np.eye(No)).conj().T
    out = _bundle(hs, st['rvec'], Ns, st['Nx'], st['Ny'], st['Nz'], st['Nw'], st['temp'], st['mu'], U, J)
    out['slist'], out['invs'] = O.soc_spin_lists(Ns)
    out['Vmat'] = O.get_vmat_soc(out['olist'], out['slist'], out['site'], out['invs'], U, J)
    return out


def full_green(st, Nw=None, mu=None):
    """G0_ab(k, slot) on every grid point (computed there, no unfolding),
    slots 0..Nw-1 = w_j, slot 2Nw-1-j = -w_j through G(k,-w) = G(k,w)^dagger:
    array [a, b, x, y, z, slot], (a, b) in the Fortran order of the library's
    Gk(k, w, a, b) (the Python array Gk[b, a, w, k])."""
    Nw = st['Nw'] if Nw is None else Nw
    mu = st['mu'] if mu is None else mu
    Nx, Ny, Nz = st['Nx'], st['Ny'], st['Nz']
    ix, iy, iz = np.meshgrid(np.arange(Nx), np.arange(Ny), np.arange(Nz), indexing='ij')
    kf = np.stack([ix.ravel() / Nx, iy.ravel() / Ny, iz.ravel() / Nz], axis=1)
    e, u = O.get_eig(O.gen_ham(kf, st['ham_r'], st['rvec']))
    G = np.swapaxes(O.gen_Green0(e, u, mu, st['temp'], Nw), 0, 1)   # Fortran (a, b): [a, b, w, k]
    Gm = np.conj(np.swapaxes(G, 0, 1))[:, :, ::-1]                  # G(k,-w) at slots Nw.. (reversed)
    full = np.concatenate([G, Gm], axis=2)                  # [a, b, slot, k]
    No = full.shape[0]
    return np.moveaxis(full, 2, 3).reshape(No, No, Nx, Ny, Nz, 2 * Nw)


def G0_at(st, Nw, mu=None):
    mu = st['mu'] if mu is None else mu
    return O.gen_Green0(st['eig'], st['uni'], mu, st['temp'], Nw)
