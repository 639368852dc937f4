"""Checkpoint and restart (fdtd_checkpoint_save / _load): a run interrupted
after n1 steps, its state saved to host memory and loaded into a fresh
handle set up the same way, continues to the same bits as the run that was
never interrupted -- fields, spectra and probe traces -- in both
precisions, both engines, across the two FUSED step forms, with dispersive
cells, PML splits, gated DFT planes, probes and a TFSF line in the state."""
import numpy as np
import pytest

from test_oracle_tfsf import tfsf_box

pytestmark = pytest.mark.gpu


def _dimer(nsteps):
    from fdtd_amd import workloads

    return workloads.c2_gold_dimer(n=64, radius=10.0, nfreq=8, nsteps=nsteps)


def _ref_box(nsteps):
    from fdtd_amd import workloads

    return workloads.ref_box_small(nsteps=nsteps)


def _tfsf(nsteps):
    spec = tfsf_box(1, 0, True, nsteps=nsteps)
    spec["tfsf_probes"] = [[0, 30]]
    return spec


def _state(sim, ids, pids):
    return {"fields": [sim.field(q) for q in range(6)], "flux": [sim.flux(i) for i in ids],
            "probes": [sim.probe(p) for p in pids]}


def _same(a, b):
    for key in ("fields", "flux", "probes"):
        for x, y in zip(a[key], b[key]):
            assert np.array_equal(x, y), key


def _run(spec, precision, n1, env_a, env_b, monkeypatch):
    import fdtd_amd

    n2 = spec["nsteps"] - n1
    for k, v in env_a.items():
        monkeypatch.setenv(k, v)
    a, ids, pids = fdtd_amd.make_simulation(spec, precision=precision)
    a.step(n1)
    ck = a.checkpoint()
    a.step(n2)
    whole = _state(a, ids, pids)
    a.close()
    for k, v in env_b.items():
        monkeypatch.setenv(k, v)
    b, ids, pids = fdtd_amd.make_simulation(spec, precision=precision)
    b.restore(ck)
    b.step(n2)
    resumed = _state(b, ids, pids)
    b.close()
    return whole, resumed, len(ck)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("forms", [("0", "0"), ("0", "1"), ("1", "0")])
def test_fused_dimer_resumes_bit_for_bit(precision, forms, monkeypatch):
    """Gold dimer, 12-cell PML on all faces, three DFT planes, 8 probes;
    saved after 150 steps (mid DFT block), resumed for 250; the saving and
    the resuming handle may use different step forms."""
    whole, resumed, _ = _run(_dimer(400), precision, 150, {"FDTD_ONEPASS": forms[0]},
                             {"FDTD_ONEPASS": forms[1]}, monkeypatch)
    _same(whole, resumed)
    assert np.abs(whole["flux"][0]).max() > 0


@pytest.mark.parametrize("xface", [("1", "0"), ("0", "1")])
def test_fp32_x_face_tiles_resume_across_settings(xface, monkeypatch):
    """FP32 one-pass step with the x slabs marched in x-face tiles or left
    to the general kernels (FDTD_XFACE): a state saved under one setting
    resumes under the other to the same bits (the split parts live in the
    same arrays either way)."""
    env = {"FDTD_ONEPASS": "1", "FDTD_ONEPASS_BY": "4", "FDTD_ONEPASS_KC": "16"}
    whole, resumed, _ = _run(_dimer(400), "fp32", 150, dict(env, FDTD_XFACE=xface[0]),
                             dict(env, FDTD_XFACE=xface[1]), monkeypatch)
    _same(whole, resumed)
    assert np.abs(whole["flux"][0]).max() > 0


@pytest.mark.parametrize("maker", [_ref_box, _tfsf])
def test_staged_and_tfsf_resume_bit_for_bit(maker, monkeypatch):
    """The reference-order engine with its y-slab PML and gated planes, and
    a TFSF line (its incident fields and PML object in the state)."""
    spec = maker(600)
    whole, resumed, _ = _run(spec, "fp64", 333, {}, {}, monkeypatch)
    _same(whole, resumed)


def test_checkpoint_refusals():
    """A handle set up differently, a truncated buffer, or a load after the
    first step are refused; nothing is loaded."""
    import fdtd_amd

    spec = _dimer(60)
    a, _, _ = fdtd_amd.make_simulation(spec, precision="fp32")
    a.step(20)
    ck = a.checkpoint()
    other = _dimer(60)
    other["grid"] = dict(other["grid"], nx=other["grid"]["nx"] + 1)
    b, _, _ = fdtd_amd.make_simulation(other, precision="fp32")
    with pytest.raises(fdtd_amd.FdtdError):
        b.restore(ck)
    c, _, _ = fdtd_amd.make_simulation(spec, precision="fp64")
    with pytest.raises(fdtd_amd.FdtdError):
        c.restore(ck)
    d, _, _ = fdtd_amd.make_simulation(spec, precision="fp32")
    with pytest.raises(fdtd_amd.FdtdError):
        d.restore(ck[:-8])
    d.step(1)
    with pytest.raises(fdtd_amd.FdtdError):
        d.restore(ck)
    for s in (a, b, c, d):
        s.close()
