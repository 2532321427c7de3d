"""Weights-only checkpoint import (tensormol_amd.tfbundle): the CRC-32C and
varint primitives against their published values, the SSTable / bundle reader
on bundles written from the documented layout (several data blocks,
prefix-compressed keys, two shards, float32 and float64), its error paths, and
the mapping of the reference's BP+EE variable names onto BPEEModel
(TFMolInstanceDirect.py:5179-5201, 5235-5264).  No checkpoint written by
TensorFlow is available offline, so agreement with TensorFlow's own files is
unpinned."""
import numpy as np
import pytest

from tensormol_amd import tfbundle as TB
from tests import tfbundle_writer as W


def test_crc32c_check_value():
    # the published check value of CRC-32C (iSCSI, RFC 3720 test vectors)
    assert TB.crc32c(b"123456789") == 0xE3069283
    assert TB.crc32c(b"\x00" * 32) == 0x8A9136AA
    assert TB.crc32c(b"\xff" * 32) == 0x62A8AB43
    assert TB.crc32c(bytes(range(32))) == 0x46DD794E


def test_varint_known_encodings():
    for v, enc in ((0, b"\x00"), (1, b"\x01"), (127, b"\x7f"), (128, b"\x80\x01"),
                   (300, b"\xac\x02"), (2 ** 32, b"\x80\x80\x80\x80\x10")):
        assert W.varint(v) == enc
        assert TB._varint(enc + b"\x05", 0) == (v, len(enc))
    with pytest.raises(ValueError):
        TB._varint(b"\x80\x80", 0)


def _tensors(rng):
    return {
        "EnergyNet/1_hidden1/weights": rng.normal(size=(6, 4)),
        "EnergyNet/1_hidden1/biaseslayer0": rng.normal(size=4).astype(np.float32),
        "EnergyNet/1_hidden1/weights/Adam": rng.normal(size=(6, 4)),
        "global_step": np.array(1234, dtype=np.int64),
        "scalar32": np.array(2.5, dtype=np.float32),
        "int_matrix": rng.integers(-5, 5, size=(3, 2)).astype(np.int32),
    }


@pytest.mark.parametrize("nshards,masked,block_entries", [(1, True, 3), (2, False, 1), (1, True, 64)])
def test_bundle_round_trip(tmp_path, nshards, masked, block_entries):
    rng = np.random.default_rng(1)
    t = _tensors(rng)
    prefix = str(tmp_path / "model-chk-10")
    W.write_bundle(prefix, t, nshards=nshards, masked_crc=masked, block_entries=block_entries)
    got = TB.read_bundle(prefix)
    assert set(got) == set(t)
    for k, v in t.items():
        assert got[k].dtype == v.dtype and got[k].shape == v.shape, k
        np.testing.assert_array_equal(got[k], v)
    sub = TB.read_bundle(prefix, names=["scalar32"])
    assert list(sub) == ["scalar32"]


def test_bundle_errors(tmp_path):
    rng = np.random.default_rng(2)
    prefix = str(tmp_path / "m")
    W.write_bundle(prefix, _tensors(rng))
    # a flipped data byte fails the tensor checksum
    path = prefix + ".data-00000-of-00001"
    raw = bytearray(open(path, "rb").read())
    raw[3] ^= 0x40
    open(path, "wb").write(bytes(raw))
    with pytest.raises(ValueError, match="checksum"):
        TB.read_bundle(prefix)
    assert len(TB.read_bundle(prefix, verify=False)) == 6
    # a flipped index byte fails the block checksum; a wrong magic is not a table
    idx = bytearray(open(prefix + ".index", "rb").read())
    idx[2] ^= 0x01
    open(prefix + ".index", "wb").write(bytes(idx))
    with pytest.raises(ValueError, match="block checksum"):
        TB.read_bundle(prefix)
    idx[2] ^= 0x01
    idx[-1] ^= 0xFF
    open(prefix + ".index", "wb").write(bytes(idx))
    with pytest.raises(ValueError, match="not a table"):
        TB.read_bundle(prefix)


def _write_bpee(prefix, model, dtype=np.float64, extra=True):
    t = {}
    nh = len(model.hidden)
    for Z in model.eles:
        for nets, names in zip((model.energy_nets, model.charge_nets),
                               TB.bpee_variable_names(Z, nh)):
            for (Wm, b), (wn, bn) in zip(nets[Z], names):
                t[wn] = Wm.astype(dtype)
                t[bn] = b.astype(dtype)
                if extra:   # optimizer slots the reader must ignore
                    t[wn + "/Adam"] = np.zeros_like(Wm, dtype=dtype)
                    t[wn + "/Adam_1"] = np.ones_like(Wm, dtype=dtype)
    if extra:
        t["beta1_power"] = np.array(0.9, dtype=dtype)
    W.write_bundle(prefix, t, block_entries=5)


def test_reference_variable_names():
    E, Q = TB.bpee_variable_names(8, 3)
    assert E[0] == ("EnergyNet/8_hidden1/weights", "EnergyNet/8_hidden1/biaseslayer0")
    assert E[2] == ("EnergyNet/8_hidden3/weights", "EnergyNet/8_hidden3/biaseslayer2")
    assert E[3] == ("EnergyNet/8_regression_linear/weights", "EnergyNet/8_regression_linear/biases")
    assert Q[1] == ("DipoleNet/8_hidden2_charge/weights", "DipoleNet/8_hidden2_charge/biases")
    assert Q[3] == ("DipoleNet/8_regression_linear_charge/weights",
                    "DipoleNet/8_regression_linear_charge/biases")


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_model_from_checkpoint(tmp_path, dtype):
    from tensormol_amd.model import BPEEModel
    src = BPEEModel([1, 8], hidden=[16, 12, 8], seed=4, init="truncnorm", dsf_alpha=0.2)
    prefix = str(tmp_path / "water-chk")
    _write_bpee(prefix, src, dtype)
    m = TB.model_from_checkpoint(prefix, [1, 8], dsf_alpha=0.2)
    assert m.hidden == [16, 12, 8] and m.eles == [1, 8]
    for Z in (1, 8):
        for a, b in ((m.energy_nets, src.energy_nets), (m.charge_nets, src.charge_nets)):
            for (Wa, ba), (Wb, bb) in zip(a[Z], b[Z]):
                np.testing.assert_array_equal(Wa, Wb.astype(dtype).astype(np.float64))
                np.testing.assert_array_equal(ba, bb.astype(dtype).astype(np.float64))


def test_model_from_checkpoint_shape_errors(tmp_path):
    from tensormol_amd.model import BPEEModel
    src = BPEEModel([1, 8], hidden=[16, 12, 8], seed=4)
    prefix = str(tmp_path / "bad")
    # drop one charge-net bias
    t_model = BPEEModel([1, 8], hidden=[16, 12, 8], seed=4)
    _write_bpee(prefix, t_model, extra=False)
    got = TB.read_bundle(prefix)
    del got["DipoleNet/8_hidden2_charge/biases"]
    W.write_bundle(prefix, got)
    with pytest.raises(KeyError, match="DipoleNet/8_hidden2_charge/biases"):
        TB.model_from_checkpoint(prefix, [1, 8])
    # a network for another descriptor width
    _write_bpee(prefix, src, extra=False)
    got = TB.read_bundle(prefix)
    for Z in (1, 8):
        got[f"EnergyNet/{Z}_hidden1/weights"] = got[f"EnergyNet/{Z}_hidden1/weights"][:-1]
    W.write_bundle(prefix, got)
    with pytest.raises(ValueError, match="inputs"):
        TB.model_from_checkpoint(prefix, [1, 8])


def test_last_checkpoint_and_elements(tmp_path):
    from tensormol_amd.model import BPEEModel
    src = BPEEModel([1, 8], hidden=[16, 12, 8], seed=4)
    d = tmp_path / "water_net"
    d.mkdir()
    for step in (9, 120, 35):
        _write_bpee(str(d / f"water_net-chk-{step}"), src, extra=False)
    (d / "water_net-chk-999.meta").write_bytes(b"")   # a graph file with no bundle
    assert TB.find_last_checkpoint(str(d), "water_net") == str(d / "water_net-chk-120")
    assert TB.find_last_checkpoint(str(tmp_path / "none"), "water_net") is None
    assert TB.checkpoint_elements(TB.read_bundle(str(d / "water_net-chk-9"))) == [1, 8]
    m = TB.model_from_checkpoint(str(d / "water_net-chk-9"))
    assert m.eles == [1, 8]


def test_manager_loads_reference_checkpoint(tmp_path):
    """TFMolManage(Name) with no .npz finds <networks_directory>/<Name>/<Name>-chk-<step>
    (TFInstance.py:94, 246-258) and takes its weights."""
    from tensormol_amd.model import BPEEModel
    from tensormol_amd.params import PARAMS
    from tensormol_amd.manage import TFMolManage
    src = BPEEModel([1, 8], hidden=[16, 12, 8], seed=7)
    d = tmp_path / "water_net"
    d.mkdir()
    _write_bpee(str(d / "water_net-chk-42"), src)
    old = PARAMS["networks_directory"]
    PARAMS["networks_directory"] = str(tmp_path)
    try:
        mgr = TFMolManage("water_net", None, False,
                          "fc_sqdiff_BP_Direct_EE_ChargeEncode_Update_vdw_DSF_elu_Normalize_Dropout",
                          False, False)
    finally:
        PARAMS["networks_directory"] = old
    assert mgr.model.eles == [1, 8] and mgr.model.hidden == [16, 12, 8]
    for Z in (1, 8):
        for (Wa, _), (Wb, _) in zip(mgr.model.energy_nets[Z], src.energy_nets[Z]):
            np.testing.assert_array_equal(Wa, Wb)
