"""KAT files computed on the GPU equal the oracle's, count for count."""
import numpy as np
import pytest

import dpcrypto_amd as D

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("scheme,n", [(D.SABER, 100), (D.FRODO976, 8)])
def test_gpu_kat_equals_oracle(oracle, scheme, n):
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from dpcrypto_amd import kat
    sd = kat.seeds(n)
    eng = D.Engine(0)
    try:
        gpu = kat.format_rsp(scheme, sd, *kat.compute(scheme, sd, eng.keypair, eng.enc, eng.dec))
    finally:
        eng.close()
    ops = (lambda sc, kc: oracle.keypair(sc, kc, threads=8), lambda sc, pk, ec: oracle.enc(sc, pk, ec, threads=8),
           lambda sc, ct, sk: oracle.dec(sc, ct, sk, threads=8))
    cpu = kat.format_rsp(scheme, sd, *kat.compute(scheme, sd, *ops))
    assert gpu == cpu
    assert len(kat.parse_rsp(gpu)) == n
