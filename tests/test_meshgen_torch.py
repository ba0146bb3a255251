"""The torch evaluation of the pressure recipe (used to build the 256^3
systems on the GPU, where the host assembly takes minutes) gives the host
recipe's T-FlowS arrays bit for bit: pattern, diagonal positions, values
and the trailing NONZ slot.  Run on CPU tensors."""
import numpy as np
import pytest


@pytest.mark.parametrize("shape", [(7, 5, 4), (16, 16, 16), (9, 12, 5), (1, 6, 3)])
def test_torch_pressure_matches_host(pkg, shape):
    mg = pkg.meshgen
    m = mg.wall_pressure_system(shape).matrix
    n, nonz, acol, arow, adia, aval = mg.pressure_matrix_torch(mg.wall_pressure_mesh(shape), device="cpu")
    assert (n, nonz) == (m.n, m.nonz)
    assert np.array_equal(acol.numpy(), m.acol) and np.array_equal(arow.numpy(), m.arow)
    assert np.array_equal(adia.numpy(), m.adia) and np.array_equal(aval.numpy(), m.aval)


def test_torch_pressure_channel(pkg):
    mg = pkg.meshgen
    mesh = mg.channel_mesh(10, 8, 6, 1.1)
    m = mg.pressure_matrix(mesh, pin_cell=1)
    _, _, acol, arow, adia, aval = mg.pressure_matrix_torch(mesh, device="cpu")
    assert np.array_equal(arow.numpy(), m.arow) and np.array_equal(aval.numpy(), m.aval)
