"""Frame descriptors of the ctypes binding: the layout the library will read.

The library reads a frame tensor's own storage through (data_ptr, frame stride,
ld), so every tensor whose storage does not have that layout must be rejected
(a silent copy would leave the library reading the original storage with the
copy's strides).  CPU only: frame_layout needs no device.
"""
import pytest
import torch

from cpptraj_amd import RmsdError, frame_layout


def test_packed_aos_layouts():
    X = torch.zeros(5, 7, 3)
    assert frame_layout(X) == (21, 0, 7)
    assert frame_layout(X.reshape(5, 21)) == (21, 0, 7)
    # padded rows (a frame stride above the row length) are read in place
    P = torch.zeros(5, 30)[:, :21]
    assert frame_layout(P) == (30, 0, 7)
    assert frame_layout(torch.zeros(5, 10, 3)[:, :7]) == (30, 0, 7)
    # natom below the row length uses the first natom atoms
    assert frame_layout(X, natom=4) == (21, 0, 4)
    # every other frame: a larger frame stride, still packed inside a frame
    assert frame_layout(X[::2]) == (42, 0, 7)
    # one frame: the stride is not read
    assert frame_layout(X[:1]) == (21, 0, 7)


def test_views_the_library_cannot_read_are_rejected():
    base = torch.zeros(6, 7, 4)
    with pytest.raises(RmsdError):
        frame_layout(base[:, :, :3])          # strides (28, 4, 1): atoms not packed
    with pytest.raises(RmsdError):
        frame_layout(torch.zeros(6, 14, 3)[:, ::2])  # every second atom
    ref = torch.zeros(1, 7, 3)
    with pytest.raises(RmsdError):
        frame_layout(ref.expand(6, 7, 3))     # stride 0: frames overlap
    with pytest.raises(RmsdError):
        frame_layout(torch.zeros(6, 21).t().contiguous().t())  # column-major rows
    with pytest.raises(RmsdError):
        frame_layout(torch.zeros(6, 7, 2))    # not xyz
    with pytest.raises(RmsdError):
        frame_layout(torch.zeros(5, 7, 3), natom=8)


def test_soa_layouts():
    S = torch.zeros(4, 3, 64)
    assert frame_layout(S, soa=True) == (192, 64, 64)
    assert frame_layout(S[:, :, :50], soa=True) == (192, 64, 50)
    assert frame_layout(S, soa=True, natom=50) == (192, 64, 50)
    with pytest.raises(RmsdError):
        frame_layout(S.transpose(1, 2), soa=True)
    with pytest.raises(RmsdError):
        frame_layout(torch.zeros(1, 3, 64).expand(4, 3, 64), soa=True)
    with pytest.raises(RmsdError):
        frame_layout(torch.zeros(4, 64, 3).permute(0, 2, 1), soa=True)
