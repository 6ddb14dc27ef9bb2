"""Qdyn6's seeded start of an MD run, on the host.

When an input gives `random_seed`, Qdyn6 draws Maxwell-Boltzmann velocities
from its own generator, satisfies the constraints for positions and
velocities, and removes the centre-of-mass motion before step 0
(src/qdyn.f90:131-135).  These functions repeat that sequence in the
reference's arithmetic order, so a device run can start from exactly the
state stock Qdyn6 would start from:

  * `QdynRandom`: randm (src/simprep.f90:4357-4391), a multiplicative
    congruential generator (multiplier 31415821, modulus 1e8) whose products
    are split into 4-digit halves so no intermediate overflows 32 bits;
    gauss (:962-977) sums 12 uniforms;
  * `maxwell` (:2491-2505): three gauss draws (x, y, z) per atom in atom
    order, standard deviation sqrt(Boltz T / m);
  * `shake` (src/bondene.f90:1069-1151) and `initial_constraint`
    (:1025-1048);
  * `stop_cm_translation` (src/simprep.f90:4496-4518).
"""
import numpy as np

BOLTZ = 0.001986  # src/globals.f90:46
DT_FS = 0.020462  # internal time units per fs (src/md.f90:207)
CONST_TOL = 0.0001  # src/globals.f90:519
CONST_MAX_ITER = 1000


class QdynRandom:
    """The reference's uniform and Gaussian generators.  Like randm, the
    state is initialised from the first seed it sees and then only advances."""

    M, M1, MULT = 100000000, 10000, 31415821

    def __init__(self, seed):
        self.irand = abs(int(seed)) % self.M

    def randm(self):
        irandh, irandl = divmod(self.irand, self.M1)
        multh, multl = divmod(self.MULT, self.M1)
        irand = ((irandh * multl + irandl * multh) % self.M1) * self.M1 + irandl * multl
        self.irand = (irand + 1) % self.M
        r = float(self.irand // 10) * 10 / float(self.M)
        if r <= 0.0 or r > 1.0:
            r = 0.0
        return r

    def gauss(self, am, sd):
        a = 0.0
        for _ in range(12):
            a = a + self.randm()
        return (a - 6.0) * sd + am


def maxwell(mass, T, seed, boltz=BOLTZ):
    """Velocities (natom, 3) as maxwell draws them: per atom in order,
    gauss(0, sd) for x, y and z with sd = sqrt(Boltz T / m)."""
    rng = QdynRandom(seed)
    kT = boltz * T
    v = np.zeros((len(mass), 3))
    for i, m in enumerate(np.asarray(mass, float)):
        sd = np.sqrt(kT / m)
        v[i, 0] = rng.gauss(0.0, sd)
        v[i, 1] = rng.gauss(0.0, sd)
        v[i, 2] = rng.gauss(0.0, sd)
    return v


def water_constraints(nat_solute, nwat, roh2, rhh2):
    """Constraint molecules of rigid 3-site waters in the order Qdyn6 builds
    them from a topology that lists O-H1, O-H2, H1-H2: (i, j) atom pairs per
    molecule and their squared lengths."""
    o = nat_solute + 3 * np.arange(nwat)
    ij = np.stack([np.stack([o, o + 1], 1), np.stack([o, o + 2], 1), np.stack([o + 1, o + 2], 1)], 1)
    return ij, np.array([roh2, roh2, rhh2])


def shake(xx, x, winv, ij, dist2):
    """shake(xx, x): moves x so each constraint (i, j) of every molecule
    holds to CONST_TOL relative, correcting along the reference bond xx(i) -
    xx(j); constraints of a molecule are visited in order, repeatedly, until
    all are in range.  ij: (nmol, nc, 2) atom indices; dist2: (nc,) or
    (nmol, nc).  Molecules are independent, so they are advanced together.
    Returns the largest iteration count."""
    nmol, nc = ij.shape[0], ij.shape[1]
    d2 = np.broadcast_to(np.asarray(dist2, float), (nmol, nc))
    ready = np.zeros((nmol, nc), bool)
    done = np.zeros(nmol, bool)
    nits = np.zeros(nmol, int)
    for _ in range(CONST_MAX_ITER):
        for c in range(nc):
            act = ~ready[:, c] & ~done
            if not act.any():
                continue
            i, j = ij[act, c, 0], ij[act, c, 1]
            vec = x[i] - x[j]
            r2 = vec[:, 0] ** 2 + vec[:, 1] ** 2 + vec[:, 2] ** 2
            diff = d2[act, c] - r2
            ready[act, c] = np.abs(diff) < CONST_TOL * d2[act, c]
            xxij = xx[i] - xx[j]
            scp = vec[:, 0] * xxij[:, 0] + vec[:, 1] * xxij[:, 1] + vec[:, 2] * xxij[:, 2]
            corr = diff / (2.0 * scp * (winv[i] + winv[j]))
            x[i] = x[i] + (xxij * corr[:, None]) * winv[i][:, None]
            x[j] = x[j] + ((-xxij) * corr[:, None]) * winv[j][:, None]
        nits[~done] += 1
        done = ready.all(1)
        if done.all():
            return int(nits.max()) if nmol else 0
    raise RuntimeError("shake failure: constraints not met in %d iterations" % CONST_MAX_ITER)


def initial_constraint(x, v, dt_int, winv, ij, dist2):
    """initial_constraint: positions first (reference = the unconstrained
    positions), then velocities through x - v dt; returns new (x, v)."""
    x = np.array(x, float)
    xx = x.copy()
    shake(xx, x, winv, ij, dist2)
    xx = x - v * dt_int
    shake(x, xx, winv, ij, dist2)
    return x, (x - xx) / dt_int


def stop_cm_translation(v, mass):
    """Removes the centre-of-mass velocity (masses summed in atom order)."""
    mass = np.asarray(mass, float)
    totmass = 0.0
    vcm = np.zeros(3)
    for i in range(len(mass)):
        totmass = totmass + mass[i]
        vcm = vcm + v[i] * mass[i]
    vcm = vcm / totmass
    return v - vcm


def seeded_start(x, mass, T, seed, dt_fs, nat_solute, nwat, roh2, rhh2):
    """Qdyn6's start for a system whose only constraints are its rigid
    waters: maxwell, initial_constraint, stop_cm_translation.  Returns
    (x, v) with v in A per internal time unit."""
    mass = np.asarray(mass, float)
    v = maxwell(mass, T, seed)
    ij, d2 = water_constraints(nat_solute, nwat, roh2, rhh2)
    x, v = initial_constraint(x, v, DT_FS * dt_fs, 1.0 / mass, ij, d2)
    return x, stop_cm_translation(v, mass)
