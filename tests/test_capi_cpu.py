"""CPU-side checks of the C ABI: the library loads, exports every symbol the
header declares, and its host-only entry points (parameter packing, class
registry, argument validation) behave like the reference's param_type
constructors, asDict() and class registration.  No kernel is launched."""

import ctypes
import os
import re

import numpy as np
import pytest

import oracle
from azplugins_amd import _azp, _azplugins

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    with open(os.path.join(ROOT, "include", "azp.h")) as fh:
        text = fh.read()
    return sorted(set(re.findall(r"^(?:int|size_t|const char\*) (azp_[a-z0-9_]+)\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    L = _azp.lib()
    names = _declared()
    assert len(names) >= 30
    for n in names:
        assert hasattr(L, n), n
    bound = {s[0] for s in _azp.SIGNATURES}
    assert set(names) == bound, set(names) ^ bound
    assert L.azp_abi_version() == _azp.ABI_VERSION == 3


KEYS = oracle.PARAM_KEYS

CASES = {
    0: [{"epsilon": 2.0}, {"epsilon": 0.0}],
    1: [{"epsilon": 1.0, "kappa": 1.0, "delta": 3.0}],
    2: [{"epsilon": 2.0, "sigma": s, "attraction_scale_factor": 0.5} for s in (1.05, 0.5, 0.85, 1.0)],
    3: [{"A": 100.0, "a_1": 1.5, "a_2": 0.0, "sigma": s} for s in (2.0, 1.05)],
    4: [{"A": 2.0, "gamma": 4.5, "s": 0.5}],
    5: [{"M_d": 1.8341, "M_r": 0.0302, "r_eq": 1.0043, "omega": 5.0, "alpha": 0.4, "repulsion": False}],
}


@pytest.mark.parametrize("ev", sorted(CASES))
def test_param_pack_matches_oracle_and_round_trips(ev):
    size = _azp.lib().azp_param_size(ev)
    for v in CASES[ev]:
        vals = [float(v[k]) for k in KEYS[ev]]
        pod = (ctypes.c_char * size)()
        _azp.call("azp_param_pack", ev, _azp.dbl_array(vals), pod)
        stored = np.frombuffer(bytes(pod)[: 8 * min(5, len(oracle.pack(ev, v)))], dtype=np.float64)
        # stored fields equal the oracle's restatement of the param_type constructor
        np.testing.assert_array_equal(stored, oracle.pack(ev, v)[: stored.shape[0]])
        back = (ctypes.c_double * 8)()
        _azp.call("azp_param_unpack", ev, pod, back)
        got = list(back)[: len(vals)]
        if ev == 5:
            assert np.allclose(got, vals)  # 1/(1/M_r) is not exact (test_pair_aniso.py:147-149)
        else:
            assert got == vals  # exact, as test_pair.py:349 requires


def test_class_registry():
    ev = ctypes.c_int()
    drv = ctypes.c_int()
    for name, e, d in [
        ("PotentialPairPerturbedLennardJonesGPU", 2, 0),
        ("PotentialPairColloid", 3, 0),
        ("PotentialPairDPDThermoGeneralWeightGPU", 4, 1),
        ("PotentialPairConservativeGeneralWeight", 4, 2),
        ("AnisoPotentialPairTwoPatchMorseGPU", 5, 3),
    ]:
        _azp.call("azp_class_lookup", name.encode(), ctypes.byref(ev), ctypes.byref(drv))
        assert (ev.value, drv.value) == (e, d)
    with pytest.raises(_azp.AzpError):
        _azp.call("azp_class_lookup", b"PotentialPairLJ", ctypes.byref(ev), ctypes.byref(drv))


def test_pair_object_host_side_validation():
    pp = _azplugins.PotentialPairDPDThermoGeneralWeightGPU(2)
    with pytest.raises(_azp.AzpError, match="mode 'none' only"):
        pp.setShiftMode("shift")
    with pytest.raises(_azp.AzpError, match="FP64 form only"):
        pp.setPrecision("mixed")
    with pytest.raises(_azp.AzpError, match="out of range"):
        pp.setParams(2, 0, [1.0, 1.0, 1.0])
    with pytest.raises(_azp.AzpError, match="not set"):
        pp.getParams(0, 1)
    pp.setParams(0, 1, [25.0, 4.5, 0.5])
    assert pp.getParams(1, 0) == [25.0, 4.5, 0.5]  # symmetric
    pa = _azplugins.AnisoPotentialPairTwoPatchMorseGPU(1)
    with pytest.raises(_azp.AzpError):
        pa.setShiftMode("xplor")
    pc = _azplugins.PotentialPairColloidGPU(1)
    with pytest.raises(_azp.AzpError, match="FP64 form only"):
        pc.setPrecision("mixed")


def test_harness_api_shape():
    """Class names, constructor signatures and keys of hoomd.azplugins.pair."""
    import inspect

    from azplugins_amd import pair

    assert sorted(pair.__all__) == ["Colloid", "DPDGeneralWeight", "ExpandedYukawa", "Hertz",
                                    "PerturbedLennardJones", "TwoPatchMorse"]
    sig = lambda c: list(inspect.signature(c.__init__).parameters)[1:]
    assert sig(pair.Hertz) == ["nlist", "default_r_cut", "default_r_on", "mode"]
    assert sig(pair.DPDGeneralWeight) == ["nlist", "kT", "default_r_cut"]
    assert sig(pair.TwoPatchMorse) == ["nlist", "default_r_cut", "mode"]
    assert pair.DPDGeneralWeight._cpp_class_name == "PotentialPairDPDThermoGeneralWeight"
    import azplugins_amd.harness as hoomd

    p = hoomd.azplugins.pair.TwoPatchMorse(nlist=hoomd.md.nlist.Cell(buffer=0.4), default_r_cut=2.0)
    with pytest.raises(KeyError):
        p.params[("A", "A")] = dict(M_d=1.0)
    p.params[("A", "A")] = dict(M_d=1, M_r=0.2, r_eq=1.0, omega=20, alpha=0.4, repulsion=True)
    assert p.params[("A", "A")]["repulsion"] is True


def _batch(sizes, boxes, ntypes=1, evaluator=5):
    h = ctypes.c_void_p()
    R = len(sizes)
    cs = (ctypes.c_uint32 * R)(*sizes)
    cb = (_azp.Box * R)(*[_azp.Box(*b) for b in boxes])
    st = _azp.lib().azp_batch_create(ctypes.byref(h), R, cs, cb, ntypes, evaluator)
    return st, h


def test_batch_host_side_validation():
    """azp_batch_* argument checks run before any device work."""
    L = _azp.lib()
    assert _batch([], [])[0] == 1
    assert _batch([10, 0], [(9.0,) * 3] * 2)[0] == 1
    assert _batch([10], [(9.0, -1.0, 9.0)])[0] == 3
    assert _batch([10], [(9.0,) * 3], evaluator=2)[0] == 4
    st, h = _batch([500, 700, 1200], [(9.41,) * 3, (10.5,) * 3, (12.6,) * 3])
    assert st == 0
    try:
        R = ctypes.c_uint32()
        N = ctypes.c_uint32()
        _azp.call("azp_batch_size", h, ctypes.byref(R), ctypes.byref(N))
        assert (R.value, N.value) == (3, 2400)
        assert L.azp_batch_set_mode(h, 2, 1, 1) == 4  # xplor: anisotropic pairs take none/shift
        assert L.azp_batch_set_mode(h, 1, 7, 1) == 1
        _azp.call("azp_batch_set_mode", h, 0, 1, 1)
        assert L.azp_batch_set_tuning(h, ctypes.byref(_azp.BatchTuning(96, 1, 4, 256, 0.0))) == 1
        assert L.azp_batch_set_tuning(h, ctypes.byref(_azp.BatchTuning(256, 0, 4, 256, 0.0))) == 1
        assert L.azp_batch_set_tuning(h, ctypes.byref(_azp.BatchTuning(256, 2, 4, 32, 0.0))) == 1
        _azp.call("azp_batch_set_tuning", h, ctypes.byref(_azp.BatchTuning(128, 2, 8, 64, 0.0)))
        t = _azp.BatchTuning()
        _azp.call("azp_batch_get_tuning", h, ctypes.byref(t))
        assert (t.block_size, t.grab, t.workgroups_per_cu, t.walk_block) == (128, 2, 8, 64)
        # compute before binding arrays / setting parameters
        assert L.azp_batch_compute(h, 1, None) == 6
        a = _azp.BatchArrays()
        for f, _ in a._fields_[:-1]:
            setattr(a, f, 0x1000)  # never dereferenced: validation fails first
        a.virial_pitch = 2432
        _azp.call("azp_batch_bind", h, ctypes.byref(a))
        assert L.azp_batch_compute(h, 1, None) == 6  # params not set
        vals = _azp.dbl_array([1.0, 0.2, 1.0, 20.0, 0.4, 1.0])
        _azp.call("azp_batch_set_params", h, 0, 0, vals)
        assert L.azp_batch_compute(h, 1, None) == 6  # r_cut not set
        _azp.call("azp_batch_set_r_cut", h, 0, 0, 2.5)
        _azp.call("azp_batch_set_r_buff", h, 2.5)  # r_list 5.0 > 9.41 / 2
        assert L.azp_batch_compute(h, 1, None) == 3
        assert b"replica 0" in L.azp_last_error()
        assert L.azp_batch_set_langevin(h, _azp.dbl_array([1.0, -1.0, 1.0]), 0.005, 1.0, 1.0, 1.0, 0) == 1
        assert L.azp_batch_set_langevin(h, _azp.dbl_array([1.0] * 3), 0.0, 1.0, 1.0, 1.0, 0) == 1
    finally:
        L.azp_batch_destroy(h)


def test_sorter_argument_checks():
    """Argument checks of the particle sort entry points (no GPU work)."""
    from azplugins_amd import _azp

    L = _azp.lib()
    h = ctypes.c_void_p()
    _azp.call("azp_sorter_create", ctypes.byref(h))
    try:
        p = _azp.Particles()
        p.N = 10
        p.box = _azp.Box(10.0, 10.0, 10.0)
        p.d_pos = 0x1000  # never dereferenced: validation fails first
        assert L.azp_sorter_order(h, ctypes.byref(p), 0, ctypes.c_void_p(0x1000), None) == 1
        assert L.azp_sorter_order(h, ctypes.byref(p), 1025, ctypes.c_void_p(0x1000), None) == 1
        assert L.azp_sorter_order(h, ctypes.byref(p), 256, None, None) == 1
        assert L.azp_sorter_order(None, ctypes.byref(p), 256, ctypes.c_void_p(0x1000), None) == 1
        p.box = _azp.Box(10.0, 0.0, 10.0)
        assert L.azp_sorter_order(h, ctypes.byref(p), 256, ctypes.c_void_p(0x1000), None) != 0
        a, b, o = (ctypes.c_void_p(x) for x in (0x1000, 0x2000, 0x3000))
        assert L.azp_gather_rows(a, b, o, 10, 12, None) == 1
        assert b"row_bytes" in L.azp_last_error()
        assert L.azp_gather_rows(a, a, o, 10, 16, None) == 1
        assert L.azp_gather_rows(a, b, None, 10, 16, None) == 1
        assert L.azp_gather_rows(None, None, None, 0, 16, None) == 0
    finally:
        L.azp_sorter_destroy(h)


def test_particle_sorter_and_trigger():
    import azplugins_amd.harness as hoomd

    t = hoomd.trigger.Periodic(5, phase=2)
    assert [s for s in range(20) if t(s)] == [2, 7, 12, 17]
    assert hoomd.trigger.as_trigger(3)(9) and not hoomd.trigger.as_trigger(3)(10)
    with pytest.raises(ValueError):
        hoomd.trigger.Periodic(0)
    sim = hoomd.Simulation(seed=1)
    (sorter,) = sim.operations.tuners
    assert isinstance(sorter, hoomd.tune.ParticleSorter)
    assert sorter.grid is None and sorter.trigger(0) and sorter.trigger(400) and not sorter.trigger(100)
    assert hoomd.tune.ParticleSorter(trigger=10, grid=64).grid == 64
    with pytest.raises(ValueError):
        hoomd.tune.ParticleSorter(grid=0)


def test_buffer_tuner_grid():
    """The buffer tuner's grid logic, driven without a GPU (timing stubbed)."""
    import azplugins_amd.harness as hoomd
    from azplugins_amd.harness.md import tune as mdtune

    class FakeNlist:
        buffer = 0.4

    nl = FakeNlist()
    t = hoomd.md.tune.NeighborListBuffer.with_grid(trigger=10, nlist=nl, maximum_buffer=0.8, minimum_buffer=0.2,
                                                   n_bins=4, n_rounds=1)
    speed = {0.2: 100.0, 0.4: 300.0, 0.6: 250.0, 0.8: 90.0, 0.3: 280.0, 0.5: 320.0}
    clock = [0.0]
    orig_sync, orig_time = mdtune.torch.cuda.synchronize, mdtune.time.perf_counter
    mdtune.torch.cuda.synchronize = lambda: None
    mdtune.time.perf_counter = lambda: clock[0]
    try:
        step = 0
        while not t.tuned and step < 1000:
            if t._current is not None and t.trigger(step):
                clock[0] += 10.0 / speed[round(t._current, 10)]
            t._act(step)
            step += 10
    finally:
        mdtune.torch.cuda.synchronize, mdtune.time.perf_counter = orig_sync, orig_time
    assert t.tuned
    assert t.best_buffer == pytest.approx(0.5)
    assert nl.buffer == pytest.approx(0.5)
    assert set(round(b, 10) for b in t.results) == {0.2, 0.4, 0.6, 0.8, 0.3, 0.5}
