"""Host side of batched solves (no GPU): merged set-ups hold every member's
arrays at the right element offsets."""
import numpy as np

from flurrypp_amd import cases
from flurrypp_amd.batch import MergedSetup, Member


def test_merged_setup_slices_match_members():
    ms = []
    for nx, ny, M in ((6, 3, 0.3), (8, 4, 0.45), (5, 3, 0.2)):
        s, U = cases.bump2d(nx=nx, ny=ny, order=2, mach=M, seed=nx)
        ms.append(Member(s, U))
    m = MergedSetup(ms)
    st = m.ele_start
    assert st[-1] == m.nEles == sum(x.setup.nEles for x in ms)
    for i, x in enumerate(ms):
        a, b = st[i], st[i + 1]
        g0, g = x.setup.geo, m.geo
        np.testing.assert_array_equal(g.pos_spts[:, a:b], g0.pos_spts)
        np.testing.assert_array_equal(g.pos_fpts[:, a:b], g0.pos_fpts)
        np.testing.assert_array_equal(g.jginv_spts[:, :, a:b], g0.jginv_spts)
        np.testing.assert_array_equal(g.detjac_spts[:, a:b], g0.detjac_spts)
        np.testing.assert_array_equal(g.norm_fpts[:, a:b], g0.norm_fpts)
        np.testing.assert_array_equal(g.da_fpts[:, a:b], g0.da_fpts)
        np.testing.assert_array_equal(m.U[:, a:b], x.U)
        np.testing.assert_array_equal(m.bnd_states[i], [x.setup.params.rhoBound, x.setup.params.uBound,
                                                        x.setup.params.vBound, 0.0, x.setup.params.pBound])
    # faces: member i's faces, element ids offset, boundary faces tagged with i
    f0 = 0
    for i, x in enumerate(ms):
        f = x.setup.faces
        n = len(f.eleL)
        np.testing.assert_array_equal(m.faces.eleL[f0:f0 + n], f.eleL + st[i])
        np.testing.assert_array_equal(m.faces.eleR[f0:f0 + n], np.where(f.eleR >= 0, f.eleR + st[i], -1))
        np.testing.assert_array_equal(m.face_bnd_state[f0:f0 + n], np.where(f.eleR >= 0, -1, i))
        f0 += n
    assert f0 == len(m.faces.eleL)
    assert m.fp.n_eles == m.nEles


def test_members_merge_only_when_every_option_but_the_free_stream_matches():
    import pytest
    from flurrypp_amd.batch import context_key
    a, Ua = cases.bump2d(nx=6, ny=3, order=2, mach=0.3, seed=1)
    b, Ub = cases.bump2d(nx=8, ny=4, order=2, mach=0.45, seed=2)      # other mesh and free stream
    c, Uc = cases.bump2d(nx=6, ny=3, order=2, mach=0.3, seed=1, cfl=0.9)  # other CFL
    d, Ud = cases.bump2d(nx=6, ny=3, order=3, mach=0.3, seed=1)        # other order
    assert context_key(a) == context_key(b)
    assert context_key(a) != context_key(c)
    assert context_key(a) != context_key(d)
    MergedSetup([Member(a, Ua), Member(b, Ub)])
    with pytest.raises(ValueError):
        MergedSetup([Member(a, Ua), Member(c, Uc)])
