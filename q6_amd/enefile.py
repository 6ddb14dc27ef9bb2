"""Qdyn6 energy files (the input of Qfep6) from the batched engine's samples.

Layout (Fortran sequential unformatted, each record framed by a 4-byte
little-endian length before and after, as Qdyn6 built with gfortran or flang
writes it):

  header (prep_sim_version, src/simprep.f90:4160-4164):
      int32 canary (1337: double precision), int32 arrays, int32 totresid,
      int32 types[arrays], int32 numres[arrays], int32 resid[totresid],
      int32 gcnum[arrays], char version[80]
  per sampled step (put_ene, src/nrgy.f90:127-158), for each state i:
      int32 i, float64 lambda, float64 total[arrays],
      float64 q[arrays][4] (bond, angle, torsion, improper),
      float64 qx[arrays][2], qq[arrays][2], qp[arrays][2], qw[arrays][2] (el, vdw),
      float64 restraint[arrays]
  then one record of the off-diagonal elements (OFFDIAG_SAVE, src/nrgy.f90:76-81):
      noffd x { int32 i, int32 j, float64 Hij, float64 rkl }

Only the full-system array is written (arrays = 1, no group contributions,
ene_header as set in src/simprep.f90:3791-3800: types(1) = NOGC (= 3, the
fourth enumerator of src/misc.f90:33-35),
numres(1) = 1, resid(:) = -1, totresid = 2).
"""
import struct

import numpy as np

from . import abi

CANARY_DOUBLE = 1337
NOGC = 3


def _record(payload):
    n = len(payload)
    return struct.pack("<i", n) + payload + struct.pack("<i", n)


def _read_records(buf):
    out, p = [], 0
    while p < len(buf):
        (n,) = struct.unpack_from("<i", buf, p)
        rec = buf[p + 4:p + 4 + n]
        (m,) = struct.unpack_from("<i", buf, p + 4 + n)
        if m != n:
            raise ValueError("corrupt record framing at byte %d" % p)
        out.append(rec)
        p += 8 + n
    return out


def frame_states(frames, nstates):
    """Per-state energy records of the batch energy history (nframes x
    QNB_FRAME_LEN): q (bond, angle, torsion, improper) of the Q bonded terms,
    qq, qp, qw (el, vdw) per state; qx = qq + qp + qw; total = q + qx
    (EQ(s)%total, src/potene.f90:229-231; the engine has no Q restraints, so
    that field is zero, as Qdyn6 writes it then)."""
    frames = np.asarray(frames, np.float64)
    nf = len(frames)
    st = np.zeros((nf, nstates, 14))
    for s in range(nstates):
        b = abi.F_Q0 + 6 * s
        qq, qp, qw = frames[:, b:b + 2], frames[:, b + 2:b + 4], frames[:, b + 4:b + 6]
        q = frames[:, abi.F_QB0 + 4 * s:abi.F_QB0 + 4 * s + 4]
        qx = qq + qp + qw
        st[:, s, 0] = q.sum(1) + qx.sum(1)     # total
        st[:, s, 1:5] = q
        st[:, s, 5:7] = qx
        st[:, s, 7:9] = qq
        st[:, s, 9:11] = qp
        st[:, s, 11:13] = qw
    return st


def frame_offd(frames, ij):
    """Per-frame off-diagonal records (i, j, Hij, rkl) from the batch energy
    history: ij[m] = (i, j) states of off-diagonal m (1-based, as in the FEP
    file), its (H, r) sampled at QNB_F_OFFD + 2m."""
    frames = np.asarray(frames, np.float64)
    return [[(i, j, float(f[abi.F_OFFD + 2 * m]), float(f[abi.F_OFFD + 2 * m + 1])) for m, (i, j) in enumerate(ij)]
            for f in frames]


def write_ene(path, lambdas, states, offd=(), offd_frames=None, version="q6_amd nonbonded engine"):
    """states: (nframes, nstates, 14) = total, q(bond, angle, torsion, improper),
    qx(el, vdw), qq(el, vdw), qp(el, vdw), qw(el, vdw), restraint.
    offd: off-diagonal records (i, j, Hij, rkl) written with every frame, or
    offd_frames: one such list per frame."""
    states = np.asarray(states, np.float64)
    nf, ns = states.shape[:2]
    lam = np.asarray(lambdas, np.float64)
    with open(path, "wb") as f:
        hdr = struct.pack("<iii", CANARY_DOUBLE, 1, 2)
        # types(1) = NOGC, numres(1) = 1, resid(1:2) = -1, gcnum(1) (not set by
        # Qdyn6 for the full system; stock Qdyn6 writes -1 there)
        hdr += struct.pack("<i", NOGC) + struct.pack("<i", 1) + struct.pack("<ii", -1, -1) + struct.pack("<i", -1)
        hdr += version.encode()[:80].ljust(80, b" ")
        f.write(_record(hdr))
        for k in range(nf):
            for s in range(ns):
                f.write(_record(struct.pack("<i", s + 1) + struct.pack("<d", lam[s]) + states[k, s].astype("<f8").tobytes()))
            rows = offd_frames[k] if offd_frames is not None else offd
            f.write(_record(b"".join(struct.pack("<iidd", int(i), int(j), float(h), float(r)) for i, j, h, r in rows)))


def read_ene(path, nstates):
    """Reads a single-array energy file: returns dict(canary, version,
    lambdas[nstates], states (nframes, nstates, 14), offd list per frame)."""
    recs = _read_records(open(path, "rb").read())
    canary, arrays, totresid = struct.unpack_from("<iii", recs[0], 0)
    if arrays != 1:
        raise ValueError("only single-array energy files are supported (arrays=%d)" % arrays)
    version = recs[0][12 + 4 * (3 * arrays + totresid):].decode(errors="replace").strip()
    body = recs[1:]
    per = nstates + 1
    if len(body) % per:
        raise ValueError("truncated energy file")
    nf = len(body) // per
    st = np.zeros((nf, nstates, 14))
    lam = np.zeros(nstates)
    offd = []
    for k in range(nf):
        for s in range(nstates):
            r = body[k * per + s]
            (idx,) = struct.unpack_from("<i", r, 0)
            if idx != s + 1:
                raise ValueError("state record out of order")
            lam[s] = struct.unpack_from("<d", r, 4)[0]
            st[k, s] = np.frombuffer(r, "<f8", 14, 12)
        o = body[k * per + nstates]
        offd.append([struct.unpack_from("<iidd", o, 24 * m) for m in range(len(o) // 24)])
    return dict(canary=canary, version=version, lambdas=lam, states=st, offd=offd)


__all__ = ["write_ene", "read_ene", "frame_states", "frame_offd"]
