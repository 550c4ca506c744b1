"""Physical pin of the oracle's far fields: the Mie series of a dielectric
cylinder.  The TE far field (ntffTE_FrequencySplit restated for one rank) has
no working reference routine to compare with, so this closed form is its pin;
the TM far field is pinned bit-exact by the golden fixtures and is checked here
too, as a control of the comparison itself.  At 25 cells per wavelength the
staircased cylinder leaves about 10 % pattern error; the other polarisation's
series, the pattern a transposed or mis-signed formula would approach, is
30-50 % away, and the mirror symmetry of normal incidence holds to 2 %."""
import pytest

from mie import compare
from oracle import pyoracle as po


@pytest.mark.parametrize("solver,pol", [(po.MPI_TM_UPML, "tm"), (po.MPI_TE_UPML, "te")])
def test_far_field_follows_mie_series(solver, pol):
    o = po.Sim(width_nm=4000, height_nm=4000, h_u_nm=20, pml=10, lambda_nm=500, angle_deg=0, stepNum=1500,
               solver=solver, model=po.MIE_CYLINDER, ntff_time=0, circle_radius_nm=500, circle_n=1.6)
    o.step(1500)
    import math
    err, other, asym = compare(o.far_field(), 2 * math.pi * 500 / 500, 1.6, pol)
    assert err < 0.15, err
    assert other > 0.3, other
    assert asym < 0.03, asym
