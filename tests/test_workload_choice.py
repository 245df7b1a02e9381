"""How config 2's i.i.d. packing differs from the benchmark's nematic-domain
packing (oracle/iid_disc_study.py), on the strict FP64 oracle: at 5,000
cells about 15% of the i.i.d. packing's phase-0 contacts are crossing rods
with the fallback normal (1, 0, 0) (ArrayMath.h:224-229), and its LCP needs
more than 10,000 BBPGD iterations (it converges within 30,000), where the
nematic-domain packing of the same size converges within a few dozen."""
import oracle
from iid_disc_study import crossing_fraction, one_phase
from pcc import generators as G


def test_iid_disc_lcp_converges_slowly_nematic_packing_fast():
    iid = G.random_disc(5000, seed=0)
    _, frac = crossing_fraction(iid, 0.05)
    assert frac > 0.1
    st, res, *_ = oracle.step(one_phase(30000), iid)
    assert st == 0 and 10000 < res["bbpgd_iterations"] < 30000
    assert res["residual"] <= 1e-6
    nem = G.dense_colony(5000, seed=0)
    _, frac = crossing_fraction(nem, 0.05)
    assert frac == 0.0
    st, res, *_ = oracle.step(one_phase(30000), nem)
    assert st == 0 and res["bbpgd_iterations"] < 100 and res["residual"] <= 1e-6
