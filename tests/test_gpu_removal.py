"""GPU re-entrainment of deposited particles by the wall shear
(TR_shear_removal, trKernelsShear.cl:13-121) through the C ABI against the
oracle.

The particle arrays and deposit counters live in device memory, as the
tracer's do; 20,000 seeded particles sit on random links (both walls, links
inside and outside the range the wall shear covers, a few indices past the
arrays), with random sizes (a few invalid), deposit timers (0, 1, 2, 3, 100),
weights and stale positions, and 15 % free and 15 % gone for good.  The call
covers particles [1000, 19000) (and, in one test, the range [1003, 18998),
whose ends are not multiples of the kernel's particle blocks): the rest must
stay as they were.  The critical shears are quantiles of the measured |tau|, so some particles stay,
some leave and some deposit for good.  Every array after the call is
bit-identical to orc_shear_removal given the same link shears, and the link
arrays kofm_removal_links reports are bit-identical to orc_removal_links; both
on the built walls and after kofm_update_walls grows a fouling layer (surface
type 1 on the links that moved 0.5 or more), with FP64 and FP32 populations.
"""
import numpy as np
import pytest

import oracle
from devmem import DeviceArray, hip
from komegafoulingmodel_amd import channel, lib
from pairs import make_pair

pytestmark = pytest.mark.gpu

KEYS = ("pos", "num_rep", "type", "dep_flag", "dep_timer", "loc", "timer")
N, OFFSET, COUNT, NSIZES = 20000, 1000, 18000, 4


def _tau_per_link(L, nBL):
    """The link shear output, entry k on tracer link bounds[0] + k (bottom)
    and bounds[2] + k - bounds[1] (top); zero elsewhere."""
    b = L.geometry("bounds")
    t = L.download("tau_bl")
    out = np.zeros(nBL)
    out[b[0]:b[0] + b[1]] = t[:b[1]]
    out[b[2]:b[2] + b[3]] = t[b[1]:b[1] + b[3]]
    return out


def _particles(nBL, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    flag = rng.integers(0, nBL, N).astype(np.int32)
    u = rng.random(N)
    flag[u < 0.15] = -1
    flag[(u >= 0.15) & (u < 0.3)] = -2
    flag[rng.integers(0, N, 20)] = nBL + 3
    typ = rng.integers(0, NSIZES, N).astype(np.int16)
    typ[rng.integers(0, N, 20)] = NSIZES
    return dict(pos=rng.uniform(0.0, 50.0, 2 * N), num_rep=rng.integers(1, 1000, N).astype(np.uint32),
                type=typ, dep_flag=flag,
                dep_timer=rng.choice(np.array([0, 1, 2, 3, 100], np.uint16), N),
                loc=rng.integers(-2, 5000, N).astype(np.int32),
                timer=rng.integers(0, 100, N).astype(np.int32))


def _constants(tau):
    a = np.abs(tau[tau != 0.0])
    q = lambda f: float(np.quantile(a, f))
    crit = np.array([q(0.2), q(0.1), q(0.4), q(0.3), q(0.6), q(0.5), q(0.8), q(0.7)])
    return dict(tau_crit=crit, drag_coeff=np.array([0.3, 0.5, 0.7, 0.9]) * 1e3,
                lift_coeff=np.array([0.2, 0.1, 0.05, 0.02]) * 1e3, mass=np.array([1.0, 2.0, 3.0, 4.0]),
                dt_wall=3.0, loc_stride=256, timer_start=1234)


def _check(h, L, w, C, seed, label, offset=OFFSET, count=COUNT):
    nBL = len(w.BL)
    links = L.removal_links(0.5)
    ref_links = oracle.removal_links(C, w.C0, w.BL, 0.5)
    for k in ("p0", "vN", "len", "surf"):
        assert np.array_equal(links[k], ref_links[k]), (label, k)
    tau = _tau_per_link(L, nBL)
    k = _constants(tau)
    host = _particles(nBL, seed)
    dev = {key: DeviceArray(h, host[key]) for key in KEYS}
    dep = DeviceArray(h, np.zeros(nBL * NSIZES, np.uint32))
    L.shear_removal({key: dev[key].addr for key in KEYS}, dep.addr, offset, count, **k,
                    foul_size_switch=0.5)
    L.synchronize()
    got = {key: dev[key].get() for key in KEYS}
    got_dep = dep.get()
    ref = {key: host[key].copy() for key in KEYS}
    ref_dep = np.zeros(nBL * NSIZES, np.uint32)
    oracle.shear_removal(ref_links, tau, ref, **k, x_min_val=w.geom_params()["x_release_pos"] - 0.5,
                         x_max_val=w.geom_params()["x_stop_pos"],
                         x_ind_start=int(np.floor(w.geom_params()["x_release_pos"] - 1.25)),
                         bl_dep=ref_dep, offset=offset, count=count)
    left = int(np.sum((host["dep_flag"] >= 0) & (ref["dep_flag"] == -1)))
    stayed = int(np.sum((ref["dep_timer"] != host["dep_timer"]) & (ref["dep_flag"] >= 0)))
    gone = int(np.sum((host["dep_flag"] >= 0) & (ref["dep_flag"] == -2)))
    surf1 = int(np.sum(links["surf"]))
    print(f"{label}: {left} re-entrained, {stayed} counting down, {gone} deposited, "
          f"{surf1} fouled-surface links, deposit total {int(ref_dep.sum())}", flush=True)
    assert left > 100 and stayed > 100 and gone > 100
    for key in KEYS:
        assert np.array_equal(got[key], ref[key], equal_nan=key == "pos"), (label, key)
        outside = np.r_[0:offset, offset + count:N]
        width = 2 if key == "pos" else 1
        g, h0 = got[key].reshape(N, width), host[key].reshape(N, width)
        assert np.array_equal(g[outside], h0[outside]), (label, key)
    assert np.array_equal(got_dep, ref_dep), label
    for d in list(dev.values()) + [dep]:
        d.free()
    return surf1


@pytest.mark.parametrize("storage", ["f64", "f32"])
def test_shear_removal_matches_oracle(storage):
    h = hip()
    w = channel.fouled_channel(256, 64, Fval=1e-5, seed=5, amp=(0.0, 4.0), width=(10.0, 60.0))
    L, _ = make_pair(w, storage)
    L.step(200, lib.OPTION_SAVE_MACRO_FIELDS)
    L.wall_shear()
    assert _check(h, L, w, w.C0, 11, f"{storage} built walls") == 0
    # a fouling layer: points 100-140 of the bottom wall raised by 0.8
    C = w.C0.copy()
    C[100:141, 1] += 0.8
    L.update_walls(C)
    L.step(20, lib.OPTION_SAVE_MACRO_FIELDS)
    L.wall_shear()
    assert _check(h, L, w, C, 12, f"{storage} fouled walls") > 30
    L.close()


def test_shear_removal_on_an_unaligned_range_matches_oracle():
    """A range that starts and ends off the kernel's blocks of 2048
    particles; the results are the same."""
    h = hip()
    w = channel.fouled_channel(256, 64, Fval=1e-5, seed=5, amp=(0.0, 4.0), width=(10.0, 60.0))
    L, _ = make_pair(w, "f32")
    L.step(200, lib.OPTION_SAVE_MACRO_FIELDS)
    L.wall_shear()
    _check(h, L, w, w.C0, 13, "f32 offset 1003", offset=1003, count=17995)
    L.close()


def test_shear_removal_argument_checks():
    w = channel.fouled_channel(256, 64, Fval=1e-5, seed=5, amp=(0.0, 4.0), width=(10.0, 60.0))
    L, _ = make_pair(w, "f32")
    k = dict(tau_crit=np.ones(2), drag_coeff=np.ones(1), lift_coeff=np.ones(1), mass=np.ones(1),
             loc_stride=256, timer_start=1)
    zero = {key: 0 for key in KEYS}
    for args in ((zero, 0, 0, 10), ({**zero, "pos": 8}, 8, -1, 10)):
        with pytest.raises(lib.KofmError) as e:
            L.shear_removal(*args, **k)
        assert e.value.code == -1
    with pytest.raises(lib.KofmError) as e:
        L.shear_removal(zero, 0, 0, 10, **{**k, "mass": np.ones(33), "drag_coeff": np.ones(33),
                                           "lift_coeff": np.ones(33), "tau_crit": np.ones(66)})
    assert e.value.code == -1
    # nothing to do: no arrays are touched
    L.shear_removal(zero, 0, 0, 0, **k)
    L.close()
