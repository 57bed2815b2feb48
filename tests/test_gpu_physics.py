"""Kernels against closed forms: zero-bias Boltzmann site occupations and the SET current."""
import numpy as np
import pytest

from kmc_dn_amd import _lib, engine
from tests import physics

pytestmark = pytest.mark.gpu


def _batch(net, K, hops, prehops, seeds0=0):
    N, P = net["N"], net["P"]
    lay = engine.Layout(N, P, net["nu"], net["kT"], net["I_0"], net["R"], net["distances"],
                        net["transitions_constant"])
    es = lay.energy_sets(net["E_constant"], net["site_energies"][N:])
    return engine.Batch(engine.make_trajs(K, lay.id, es[0], np.arange(seeds0, seeds0 + K), hops,
                                          net["occupation"], prehops=prehops))


@pytest.mark.parametrize("case", ["grand-canonical", "canonical"])
def test_boltzmann_site_occupations_fp32(gpu_engine, case):
    if case == "grand-canonical":
        net = physics.boltzmann_network()
        exact = physics.boltzmann_site_occupation(net["distances"], net["E_constant"],
                                                  net["I_0"], net["R"], V_e=net["V_e"])
    else:   # validation/boltzmann_validation.py: N=10, no electrodes, 5 carriers
        net = physics.boltzmann_network(seed=5, N=10, P=0, I_0=4.0, a=0.4, M=0)
        net["occupation"] = np.array([1] * 5 + [0] * 5, bool)
        exact = physics.boltzmann_site_occupation(net["distances"], net["E_constant"],
                                                  net["I_0"], net["R"], n_fixed=5)
    N = net["N"]
    b = _batch(net, 512, 40000, 2000)
    b.run(precision=_lib.FP32_BATCHED, record=True, post_hop_record=False)
    b.sync()
    r = b.results()
    assert (r.status == 0).all()
    occ = r.occ_time[:, :N] / r.time[:, None]
    m = occ.mean(0)
    se = occ.std(0, ddof=1) / np.sqrt(len(occ))
    print(f"[gpu] boltzmann {case}: max |sim-exact| {np.abs(m - exact).max():.4f}, "
          f"max z {np.max(np.abs(m - exact) / se):.2f}", flush=True)
    assert np.all(np.abs(m - exact) < 4 * se + 2e-3), (m, exact, se)


@pytest.mark.parametrize("precision", [_lib.FP32_BATCHED, _lib.FP64_DENSE])
def test_set_current(gpu_engine, precision):
    for E_G, V_S in [(0.0, 2.0), (1.0, -1.5), (-0.5, 0.7), (3.0, 5.0)]:
        net = physics.set_network(E_G, V_S)
        exact = physics.set_current(E_G, V_S, 0.0, net["transitions_constant"][1, 0],
                                    net["transitions_constant"][0, 2])
        b = _batch(net, 256, 20000, 0)
        b.run(precision=precision)
        b.sync()
        r = b.results()
        cur = r.eo[:, 1] / r.time
        se = cur.std(ddof=1) / np.sqrt(len(cur))
        print(f"[gpu] SET E_G={E_G} V_S={V_S}: {cur.mean():.5f} vs exact {exact:.5f} "
              f"(se {se:.1e})", flush=True)
        assert abs(cur.mean() - exact) < 4 * se + 2e-3 * abs(exact)
