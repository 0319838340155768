"""Configuration canonicalisation (Configuration.java:53-198,243-297).

CPU: the oracle (oracle/canon_ref.py) against the golden vectors of the
reference's ConfigurationTest.java (signatures, all 2n equivalent signatures,
canonical signatures, hash codes of three configurations), the signature
round trip through ofSignature, and the closed property that every
configuration equivalent under rotation or mirroring has the same canonical
form; argument checks of the C ABI without a GPU.  GPU: the batched kernel
(tdp_canonical_batch) bit-exact against the oracle on the golden
configurations and on random ones (arbitrary pi, up to n = 64).
"""
import json
import os
import random

import numpy as np
import pytest

import tdp1375_amd as T
from oracle import canon_ref as C
from tdp1375_amd import canonical as K

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "configuration_vectors.json")


def golden():
    with open(GOLDEN) as fh:
        return json.load(fh)["configurations"]


def f32(xs):
    return np.asarray(xs, np.float32)


def random_config(rng, n, shuffle_pi):
    while True:
        p = list(range(n))
        rng.shuffle(p)
        if all(p[x] != x for x in range(n)):
            break
    pi = list(range(n))
    if shuffle_pi:
        rng.shuffle(pi)
    return p, pi


# ---------------------------------------------------------------- CPU
def test_oracle_matches_configuration_test_vectors():
    for g in golden():
        cyc = C.parse_cycles(g["spi"])
        n = sum(len(c) for c in cyc)
        pi = list(range(n))
        assert np.array_equal(C.signature(cyc, pi), f32(g["signature"])), g["spi"]
        got = sorted(tuple(s.tolist()) for s, _, _ in C.equivalent_signatures(cyc, pi))
        want = sorted(tuple(f32(s).tolist()) for s in g["equivalent"])
        assert got == want, g["spi"]
        sig, h, _ = C.canonical(cyc, pi)
        assert np.array_equal(sig, f32(g["canonical"])), g["spi"]
        assert C.config_hash(cyc, pi) == g["hash"] == h, g["spi"]
        # ConfigurationTest round trip: ofSignature(signature).signature
        assert np.array_equal(C.signature(C.of_signature(C.signature(cyc, pi)), pi), C.signature(cyc, pi))


def test_equivalent_configurations_share_the_canonical_form():
    """Rotations and mirrors act as a group on the signatures, so the
    configuration behind any equivalent signature has the same canonical
    signature and hash (pins the least-hash choice and ofSignature)."""
    rng = random.Random(7)
    for it in range(60):
        n = rng.randint(3, 16)
        p, pi = random_config(rng, n, it % 2 == 1)
        cyc = C.cycles_of(p)
        sig, h, _ = C.canonical(cyc, pi)
        for eq, _, _ in rng.sample(C.equivalent_signatures(cyc, pi), 4):
            other = C.of_signature(eq)
            assert np.array_equal(C.signature(other, list(range(n))), eq)
            s2, h2, _ = C.canonical(other, list(range(n)))
            assert h2 == h and np.array_equal(s2, sig)


def test_argument_errors_without_gpu():
    with pytest.raises(T.TdpError, match="fixed point"):
        K.canonical_batch([[1, 0, 2]])
    with pytest.raises(T.TdpError, match="permutations"):
        K.canonical_batch([[1, 0, 0]])
    with pytest.raises(T.TdpError, match="1..64"):
        K.canonical_batch([list(range(1, 65)) + [0]])


# ---------------------------------------------------------------- GPU
@pytest.mark.gpu
def test_kernel_matches_configuration_test_vectors(gpu_lib):
    gs = golden()
    out = K.canonical_batch([C.image_of(C.parse_cycles(g["spi"]), len(g["signature"])) for g in gs])
    for g, r in zip(gs, out):
        assert np.array_equal(r.signature, f32(g["canonical"])), g["spi"]
        assert r.hash == g["hash"], g["spi"]
        assert C.cycles_of(r.spi) == C.cycles_of(C.image_of(C.of_signature(f32(g["canonical"])), len(r.spi)))
    one = K.canonical("(0 9 7)(1 5 2 6 3)(4 10 8)")
    assert one.hash == -2127296534


@pytest.mark.gpu
def test_kernel_matches_oracle_on_random_configurations(gpu_lib):
    rng = random.Random(11)
    spis, pis, want = [], [], []
    for it in range(400):
        n = rng.choice([2, 3, 5, 8, 13, 21, 34, 47, 64]) if it % 4 == 0 else rng.randint(2, 40)
        p, pi = random_config(rng, n, it % 2 == 1)
        spis.append(p)
        pis.append(pi)
        want.append(C.canonical(C.cycles_of(p), pi))
    got = K.canonical_batch(spis, pis)
    for it, (g, (sig, h, idx)) in enumerate(zip(got, want)):
        assert np.array_equal(g.signature, sig), it
        assert g.hash == h, it
        eqs = C.equivalent_signatures(C.cycles_of(spis[it]), pis[it])
        assert np.array_equal(eqs[g.index][0], sig), it   # the index names an equivalent with that content
        assert C.cycles_of(g.spi) == C.cycles_of(C.image_of(C.of_signature(sig), len(sig))), it
