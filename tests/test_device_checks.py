"""The blocked CLOUDSC2 / CLOUDST / CLOUD_S_LAYER wrappers check every device tensor before the launch: dtype of
the precision, blocked shape (KLEV+1 for half levels), contiguity, device, and a CETA of KLEV values.  A wrong
tensor raises TypeError and no kernel runs."""
import pytest

torch = pytest.importorskip("torch")

from openifs_amd import cloudsc as C
from openifs_amd import cloudsc2 as C2

NPROMA, KLEV, NGPBLKS = 64, 137, 2


def _fields(spec, device, dtype, shape=C2.blocked_shape):
    return {e[0]: torch.zeros(shape(e[1], NPROMA, KLEV, NGPBLKS), dtype=dtype, device=device) for e in spec}


def _launch(which, fields, ceta, prec):
    p = C.default_params(900.0, KLEV)
    p2 = C2.default_params2(p)
    fn = {"cloudsc2": C2.cloudsc2_blocked, "cloudst": C2.cloudst_blocked, "layer": C2.cloud_s_layer_blocked}[which]
    fn(fields, p, p2, ceta, NPROMA * NGPBLKS, NPROMA, KLEV, 900.0, precision=prec)


SPECS = {"cloudsc2": (C2.CLOUDSC2_ARRAYS, C2.blocked_shape), "cloudst": (C2.CLOUDST_ARRAYS, C2.blocked_shape),
         "layer": (C2.LAYER_ARRAYS, C2.layer_shape)}


@pytest.mark.parametrize("which", sorted(SPECS))
def test_host_tensors_are_rejected(which):
    spec, shape = SPECS[which]
    f = _fields(spec, "cpu", torch.float32, shape)
    with pytest.raises(TypeError, match="CUDA"):
        _launch(which, f, torch.zeros(KLEV), "sp")


@pytest.mark.parametrize("which", ["cloudsc2", "cloudst"])
def test_missing_array_is_rejected(which):
    spec, shape = SPECS[which]
    f = _fields(spec, "cpu", torch.float32, shape)
    del f["pfplsl"]
    with pytest.raises(KeyError, match="pfplsl"):
        _launch(which, f, torch.zeros(KLEV), "sp")


@pytest.mark.gpu
@pytest.mark.parametrize("which", sorted(SPECS))
def test_wrong_device_tensors_are_rejected_before_launch(which):
    spec, shape = SPECS[which]
    dev = "cuda"
    ceta = torch.zeros(KLEV, dtype=torch.float32, device=dev)
    # an FP32 tensor given to the FP64 entry
    f = _fields(spec, dev, torch.float32, shape)
    with pytest.raises(TypeError, match="float64"):
        _launch(which, f, ceta.double(), "dp")
    # a full-level shape where the half-level one is expected
    f = _fields(spec, dev, torch.float32, shape)
    f["pfplsl"] = torch.zeros((NGPBLKS, KLEV, NPROMA), device=dev)
    with pytest.raises(TypeError, match="pfplsl"):
        _launch(which, f, ceta, "sp")
    # a non-contiguous view
    f = _fields(spec, dev, torch.float32, shape)
    f["pfplsn"] = torch.zeros((NGPBLKS, NPROMA, KLEV + 1), device=dev).transpose(1, 2)
    with pytest.raises(TypeError, match="pfplsn"):
        _launch(which, f, ceta, "sp")
    # a short CETA
    f = _fields(spec, dev, torch.float32, shape)
    with pytest.raises(TypeError, match="ceta"):
        _launch(which, f, ceta[:-1], "sp")
    torch.cuda.synchronize()
