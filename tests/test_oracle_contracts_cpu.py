"""The CPU oracle against the reference's raycast, integration-band and marching-cubes
contracts (hand-derived expectations; see tests/contracts.py)."""
import numpy as np

import oracle
from tests import contracts as C

I4 = np.eye(4, dtype=np.float32)


def _cast(t, w, lo, hi, K=C.RC_K, W=C.RC_IMG, H=C.RC_IMG):
    v = oracle.Volume(C.RC_RES, C.RC_VS, C.RC_TRUNC, 128.0, C.RC_ORIGIN, lo, hi)
    v.set(t, w)
    return v.raycast(I4, K, W, H)


def test_raycast_contract_oracle():
    C.check_raycast_contract(_cast)


def test_thin_feature_contract_oracle():
    def cast1(t, w, lo, hi):
        v, n = _cast(t, w, lo, hi, K=C.TF_K, W=1, H=1)
        return v[0, 0], n[0, 0]

    C.check_thin_feature_contract(cast1)


def test_integrate_band_contract_oracle():
    p = C.BAND_PARAMS

    def integrate(d, lo, hi):
        v = oracle.Volume(p["resolution"], p["voxel_size"], p["truncation"], p["max_weight"], p["origin"])
        v.integrate(d, C.BAND_K, I4, lo, hi)
        return v.get()

    C.check_band_contract(integrate)


def test_mc_contracts_oracle():
    def extract(t, w, vs):
        return oracle.mc_extract(t, w, vs, (0.0, 0.0, 0.0))

    C.check_mc_contracts(extract)
