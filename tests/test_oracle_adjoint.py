"""The FP64 CPU reference of the energy E = sum_t (1 - q_M(t))^2 and of its
gradient (oracle/trinity_adjoint.c), pinned by mathematics:

  * E equals the sum of (1 - q)^2 over the reference-pinned element
    qualities of the reference-pinned metric (orc_metric_run,
    orc_quality_cache), bit for bit;
  * the gradient equals central finite differences (step 1e-6) of E, to
    1e-6 of the largest gradient entry, in x, y and f, for p = 1, 2, 4, on
    fields whose Hessians stay away from the metric's non-smooth points
    (eigenvalue clamp, the eigen solver's diagonal shortcut);
  * invariants: E is unchanged by a translation of the mesh, so the
    coordinate gradient sums to zero; E is unchanged by adding a constant to
    f, so the nodal gradient sums to zero.
"""
import numpy as np
import pytest

import oracle
from trinity_amd import meshgen as mg


def _smooth_field(m):
    x, y = m.xy[:, 0], m.xy[:, 1]
    return np.sin(3 * x) * np.cos(2 * y) + 2 * x * x + 0.5 * x * y + 3 * y * y


def _meshes():
    a = mg.grid_mesh(10, 0.25, 3)
    a.solut = _smooth_field(a)
    b = mg.workload("jitter14:fronts+quad", seed=5)
    return [a, b]


def test_energy_is_sum_of_reference_qualities():
    for m in _meshes():
        st = oracle.metric_stages(m.xy, m.tri, m.solut, 0.8, 2)
        q = oracle.quality_cache(m.xy, m.tri, st["metric"])
        E, ten, qa = oracle.energy(m.xy, m.tri, m.solut, 0.8, 2, topo=st["topo"], want_fields=True)
        assert np.array_equal(ten, st["metric"])
        assert np.array_equal(qa, q)
        assert E == pytest.approx(float(np.sum((1 - q) ** 2)), rel=1e-13)


@pytest.mark.parametrize("p", [1, 2, 4])
def test_gradient_matches_central_differences(p):
    h = 1e-6
    for m in _meshes():
        topo = oracle.Topology(m.nv, m.tri)
        E, gx, gf = oracle.energy_grad(m.xy, m.tri, m.solut, 0.8, p, topo=topo)
        assert E == oracle.energy(m.xy, m.tri, m.solut, 0.8, p, topo=topo)
        scale_x, scale_f = np.abs(gx).max(), np.abs(gf).max()
        rng = np.random.default_rng(11)
        for i in rng.choice(m.nv, 8, replace=False):
            for c in range(2):
                xp, xm = m.xy.copy(), m.xy.copy()
                xp[i, c] += h
                xm[i, c] -= h
                fd = (oracle.energy(xp, m.tri, m.solut, 0.8, p, topo=topo)
                      - oracle.energy(xm, m.tri, m.solut, 0.8, p, topo=topo)) / (2 * h)
                assert abs(fd - gx[i, c]) <= 1e-6 * scale_x, (p, i, c, fd, gx[i, c])
            fp, fm = m.solut.copy(), m.solut.copy()
            fp[i] += h
            fm[i] -= h
            fd = (oracle.energy(m.xy, m.tri, fp, 0.8, p, topo=topo)
                  - oracle.energy(m.xy, m.tri, fm, 0.8, p, topo=topo)) / (2 * h)
            assert abs(fd - gf[i]) <= 1e-6 * scale_f, (p, i, fd, gf[i])


def test_gradient_invariants():
    for m in _meshes():
        E, gx, gf = oracle.energy_grad(m.xy, m.tri, m.solut, 0.8, 2)
        # translation invariance of E: sum of coordinate gradients vanishes
        assert abs(gx[:, 0].sum()) <= 1e-9 * np.abs(gx).sum()
        assert abs(gx[:, 1].sum()) <= 1e-9 * np.abs(gx).sum()
        # f enters through its derivatives only: constants do not matter
        assert abs(gf.sum()) <= 1e-9 * np.abs(gf).sum()
        E2 = oracle.energy(m.xy + np.array([0.25, -0.5]), m.tri, m.solut + 3.0, 0.8, 2)
        assert E2 == pytest.approx(E, rel=1e-9)
