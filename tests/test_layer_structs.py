"""layer_structs (the ytf_layer array every C-ABI call takes) against a
field-by-field fill of the ctypes structs: eps and mu column-major as
interleaved (re, im), thickness, material kinds; byte-identical on every
foundation stack and the benchmark stacks (CPU only)."""
import numpy as np

from helpers import FOUNDATION_STACKS, c1_stack, c2_stack


def _fieldwise(stack):
    from yellowthinfilms_amd import _lib
    from yellowthinfilms_amd.materials import stack_arrays
    eps, mu, d, ek, mk = stack_arrays(stack)
    arr = (_lib.YtfLayer * len(d))()
    for l in range(len(d)):
        e = eps[l].T.reshape(-1)
        m = mu[l].T.reshape(-1)
        for i in range(9):
            arr[l].eps[2 * i], arr[l].eps[2 * i + 1] = e[i].real, e[i].imag
            arr[l].mu[2 * i], arr[l].mu[2 * i + 1] = m[i].real, m[i].imag
        arr[l].d = d[l]
        arr[l].eps_kind = int(ek[l])
        arr[l].mu_kind = int(mk[l])
    return arr


def test_layer_structs_match_a_fieldwise_fill():
    from yellowthinfilms_amd import workloads as W
    from yellowthinfilms_amd.green import layer_structs
    stacks = list(FOUNDATION_STACKS.values()) + [c1_stack(), c2_stack(), W.gyro_512_stack()]
    for st in stacks:
        a = layer_structs(st)
        assert len(a) == len(st)
        assert bytes(a) == bytes(_fieldwise(st))
        # spot check: layer 0's eps[0][1] (column-major entry 3) lands in eps[6], eps[7]
        e = np.stack([l.eps.eps for l in st])
        assert a[0].eps[6] == e[0][0, 1].real and a[0].eps[7] == e[0][0, 1].imag
