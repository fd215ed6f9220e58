"""The seeded input generator: SplitMix64's published first outputs and the
shapes / ranges of the synthetic images."""
import numpy as np

from tucuda_amd import synth


def test_splitmix64_known_values():
    # First outputs of SplitMix64 seeded with 0 and 1234567 (the published
    # reference implementation, Steele, Lea and Flood 2014).
    assert [hex(int(v)) for v in synth.splitmix64(0, 3)] == \
        ["0xe220a8397b1dcdaf", "0x6e789e6aa1b965f4", "0x6c45d188009454f"]
    assert hex(int(synth.splitmix64(1234567, 1)[0])) == "0x599ed017fb08fc85"


def test_images():
    a = synth.uniform_image(0, 480, 640)
    assert a.dtype == np.float32 and a.shape == (480, 640) and 0 <= a.min() and a.max() < 255
    assert np.array_equal(a, synth.uniform_image(0, 480, 640))
    e = synth.edge_image(1, 108, 192)
    assert e.max() <= 255 + 8 * 64 and e.min() >= 0
    u = synth.u8_image(2, 40, 50)
    assert u.dtype == np.uint8 and u.max() > 250 and u.min() < 5
    n = synth.normal(3, 20000)
    assert abs(n.mean()) < 0.03 and abs(n.std() - 1) < 0.03
