"""Generator / add_initial_model host logic (src/NN.jl:40-158, src/Utils.jl:140-166) on CPU."""
import math

import pytest
import torch


def test_generator_shapes_and_range(ads):
    torch.manual_seed(0)
    # NNFWI.jl:60-62 on the Marmousi2 grid (202 x 68): num_layer 4, w0 = ceil(202/16), h0 = ceil(68/16)
    G = ads.Generator(num_layer=4, w0=math.ceil(202 / 16), h0=math.ceil(68 / 16), vmin=-1.5, vmax=1.5)
    z = torch.rand(1, 8)
    out = G(z)
    assert out.shape == (13 * 16, 5 * 16)
    assert float(out.detach().abs().max()) <= 1.5
    # filters 2^(num_layer-i+4): 128, 64, 32, 16, then 1
    assert [c.out_channels for c in G.convs] == [128, 64, 32, 16] and G.out.out_channels == 1
    assert all(c.bias is None for c in G.convs) and G.dense.bias is None
    two = G(torch.rand(2, 8))
    assert two.shape == (2, 208, 80)


def test_generator_same_padding_matches_keras_convention(ads):
    """Keras 'same' for an even 4x4 kernel pads 1 before and 2 after: a delta image
    convolved with a kernel of one non-zero tap lands where that convention puts it."""
    x = torch.zeros(1, 1, 6, 6)
    x[0, 0, 2, 3] = 1.0
    w = torch.zeros(1, 1, 4, 4)
    w[0, 0, 0, 0] = 1.0  # output[i, j] = xpad[i, j] = x[i - 1, j - 1]
    y = torch.nn.functional.conv2d(ads.Generator._same(x), w)
    assert y.shape == (1, 1, 6, 6)
    assert y[0, 0, 3, 4] == 1.0 and float(y.sum()) == 1.0


def test_add_initial_model_crop_and_embed(ads):
    v = torch.zeros(202, 68)
    big = torch.arange(208 * 80, dtype=torch.float32).reshape(208, 80)
    out = ads.add_initial_model(big, v)  # crop around the centre (Utils.jl:155-160)
    assert torch.equal(out, big[3:205, 6:74])
    small = torch.ones(200, 60)
    out = ads.add_initial_model(small, v)  # add into the middle (Utils.jl:146-153)
    assert float(out.sum()) == 200 * 60 and float(out[1:201, 4:64].min()) == 1.0
    assert float(out[0].abs().sum()) == 0 and float(out[:, :4].abs().sum()) == 0
    batch = ads.add_initial_model(torch.ones(3, 202, 68), v)
    assert batch.shape == (3, 202, 68)
    with pytest.raises(ValueError):
        ads.add_initial_model(torch.ones(210, 60), v)
