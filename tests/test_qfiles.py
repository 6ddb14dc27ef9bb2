"""The synthetic workloads as stock Qdyn6 sees them.

q6_amd.qfiles writes a synthetic system as Q topology, FEP and MD input files.
Stock Qdyn6 modules built from the reference sources (oracle/_ref/qnb_dump,
made by `make -C oracle/ref` in the development container) read them and dump
their set-up and every nonbonded term.  Three checks per workload:

  * every table Qdyn6 derives (scaled charges, types and LJ codes, charge
    groups and switch atoms, exclusion and 1-4 windows, the water and
    Q-atom precompute tables, Q-atom connectivity) equals the generator's
    own, so the benchmark's systems are systems Qdyn6 builds;
  * the oracle, given the generator's tables and Qdyn6's coordinates (the
    text round trip can move a coordinate by an ulp), reproduces Qdyn6's
    energy and gradient of every term bit for bit at the workload's size:
    3,387 atoms (20 A sphere), the benchmarked 10,861 atoms (30 A), and an
    18 A sphere of a nine-site flexible solvent (general solvent loops and
    nonbond_solvent_internal, set up without the MD step);
  * the Fortran fill module (q6_amd/fortran/qnb_fill.f90), linked into the
    dump, fills the engine's structs from that Qdyn6 state with the same
    bytes the Python host passes.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

from conftest import GOLDEN, ROOT
from q6_amd import abi, qfiles, synth
from q6_amd.system import from_golden

DUMP = os.path.join(ROOT, "oracle", "_ref", "qnb_dump")
pytestmark = pytest.mark.skipif(not os.path.exists(DUMP), reason="reference build oracle/_ref/qnb_dump absent")


@pytest.fixture(scope="module", params=["cfg1", "cfg2", "cfg_ethanol"])
def written(request, tmp_path_factory):
    sys.path.insert(0, GOLDEN)
    from make_golden import compact
    from qdump import read_dump
    s = getattr(synth, request.param)()
    d = str(tmp_path_factory.mktemp(request.param))
    inp = qfiles.write_case(s, d, 10, name=request.param, seed=0)
    # the flexible-solvent system is set up without its MD step
    nomd = ["nomd"] if "solvent" in s.md else []
    r = subprocess.run([DUMP, "out.bin"] + nomd + [os.path.basename(inp)], cwd=d, capture_output=True, text=True)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    np.savez(os.path.join(d, "dump.npz"), **compact(read_dump(os.path.join(d, "out.bin"))))
    g = from_golden(os.path.join(d, "dump.npz"))
    from oracle.oracle import Oracle
    o = Oracle(s, ref_quirks=1)
    o.make_pair_lists(g.x)
    return s, g, o


def test_tables_equal_generator(written):
    s, g, _ = written
    for k in ("natom", "nat_solute", "nwat", "ncgp", "ncgp_solute", "nqat", "nstates", "ntypes", "ivdw_rule"):
        assert s.n[k] == g.n[k], k
    assert set(s.a) == set(g.a)
    for k in s.a:
        assert np.array_equal(np.asarray(s.a[k]), np.asarray(g.a[k])), k
    for k in ("Rcpp", "Rcpw", "Rcww", "Rcq", "RcLRF", "rexcl_o", "rwat", "fk_wsphere", "fkwpol", "lambda",
              "xpcent", "xwcent", "use_LRF", "wpol_restr", "nwpolr_shell", "Dwmz", "awmz", "wshell_rout", "wshell_dr",
              "wshell_cstb"):
        if k.startswith("wshell") and not s.par["wpol_restr"]:
            assert k not in g.par, k  # no polarisation shells without the restraint
            continue
        assert np.array_equal(np.asarray(s.par[k], float), np.asarray(g.par[k], float)), k
    assert int(s.n.get("n_solv_int", 0)) == int(g.n.get("n_solv_int", 0))
    assert np.max(np.abs(s.x - g.x)) <= 1e-14 * np.max(np.abs(s.x))


def test_fortran_fill_equals_host_tables(written):
    """q6_amd/fortran/qnb_fill.f90, run inside the same Qdyn6 set-up (the
    dump links it and reads every array back through its C pointer), fills
    qnb_system and qnb_params with exactly the arrays and values the Python
    host passes to qnb_create for the same system: same bytes, same layout."""
    s, g, _ = written
    r = g.ref
    for k in ("natom", "nat_solute", "nwat", "solv_atom", "ncgp", "ncgp_solute", "nqat", "nstates", "ntypes",
              "ivdw_rule", "solvent_spc"):
        assert int(r["fill_" + k]) == int(s.n[k]), k
    assert int(r["fill_any_atom_cutoff"]) == int(s.n.get("any_atom_cutoff", 0))
    assert int(r["fill_nexlong"]) == int(s.n.get("nexlong", 0))
    assert int(r["fill_n14long"]) == int(s.n.get("n14long", 0))
    assert float(r["fill_el14_scale"]) == float(s.el14_scale)
    sysc, parc = s.to_structs()
    for k, dt in s.ARRAYS.items():
        want = np.asarray(s.a[k]).ravel()
        if "fill_" + k not in r:
            assert want.size == 1 and not want.any(), k  # a table the system does not have
            continue
        got = np.asarray(r["fill_" + k]).ravel().astype(dt)
        assert got.dtype == want.dtype and np.array_equal(got, want), k
    assert np.array_equal(r["fill_cutoffs"], [parc.Rcpp, parc.Rcpw, parc.Rcww, parc.Rcq, parc.RcLRF])
    assert list(r["fill_flags"]) == [parc.use_LRF, abi.QNB_PREC_FP64, 1, parc.wpol_restr, parc.nwpolr_shell]
    assert np.array_equal(r["fill_xpcent"], list(parc.xpcent)) and np.array_equal(r["fill_xwcent"], list(parc.xwcent))
    assert np.array_equal(r["fill_lambda"], list(parc.lambda_))
    assert np.array_equal(r["fill_scaas"], [parc.rexcl_o, parc.rwat, parc.fk_wsphere, parc.Dwmz, parc.awmz, parc.Boltz,
                                            parc.fkwpol, parc.itdis])
    wsh = np.asarray(r["fill_wshell"]).T  # (shell, field) as dumped
    for row, nm in zip(wsh, ("wshell_rout", "wshell_dr", "wshell_cstb", "wshell_theta_corr")):
        assert np.array_equal(row, list(getattr(parc, nm))), nm


@pytest.mark.parametrize("term,mask", [("ww", abi.T_WW), ("qw", abi.T_QW), ("pw", abi.T_PW), ("pp", abi.T_PP),
                                       ("qp", abi.T_QP), ("lrf", abi.T_LRF), ("qq", abi.T_QQ), ("qqp", abi.T_QQP),
                                       ("solvent_radial", abi.T_RAD), ("water_pol", abi.T_POL),
                                       ("solv_int", abi.T_SINT)])
def test_oracle_matches_qdyn6_on_workload(written, term, mask):
    from test_oracle_golden import term_energy
    s, g, o = written
    if "E_" + term not in g.ref:
        pytest.skip("term not in this system")
    d, E, EQ = o.eval_terms(mask, g.x, Tfree=g.Tfree, istep=0)
    assert np.array_equal(d, g.ref["d_" + term].T)
    assert np.array_equal(np.array(term_energy(term, E, EQ)), np.atleast_1d(g.ref["E_" + term]))
