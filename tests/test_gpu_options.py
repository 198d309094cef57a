"""Non-default Poisson options on the GPU against the CPU oracle, bit for
bit: the Wiener stage matching on the observation instead of the pilot
(poisson_wiener_match_source, bmndalgo.py:714-720; its counts then use the
count-table matcher) and taking the observation's variance model
(poisson_variance_source_wiener, bmndalgo.py:635-646), alone and together."""

import pytest

import oracle
from bmnd_amd import BM3DProfile, BM4DProfile, NoiseModel, phantoms
from bmnd_amd.enums import PoissonVarianceSource

from test_gpu_parity import _compare

pytestmark = pytest.mark.gpu

OBS, PILOT = PoissonVarianceSource.OBSERVATION, PoissonVarianceSource.PILOT


@pytest.mark.parametrize("match_src,var_src", [(OBS, PILOT), (PILOT, OBS), (OBS, OBS)])
@pytest.mark.parametrize("kind", ["volume", "image"])
def test_wiener_sources_match_oracle(kind, match_src, var_src):
    print(f"options: {kind}, match {match_src.value}, variance {var_src.value}", flush=True)
    if kind == "volume":
        _, noisy = phantoms.ellipsoid_volume((26, 28, 24), seed=71, peak=10.0, n_ellipsoids=6,
                                             axes_range=(3.0, 10.0))
        prof = BM4DProfile(noise_model=NoiseModel.POISSON)
    else:
        _, noisy = phantoms.blob_image(72, seed=72, peak=20.0)
        prof = BM3DProfile(noise_model=NoiseModel.POISSON)
    prof.poisson_wiener_match_source = match_src
    prof.poisson_variance_source_wiener = var_src
    params = oracle.default_params(noisy.ndim, noisy.shape)
    params.wi_match_observation = int(match_src is OBS)
    params.wi_var_observation = int(var_src is OBS)
    stats, _, _ = _compare(noisy, prof, params, "auto-fast")
    assert stats["ht_groups_differing"] == 0
    assert stats["wiener_groups_differing"] == 0
    assert stats["pilot_bits_differing"] == 0
    assert stats["out_bits_differing"] == 0
    if match_src is OBS:
        assert stats["modes"]["wiener"] == 2  # integer counts: the count table
