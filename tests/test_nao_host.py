"""Host-side logic of neural AO (no GPU): parameter layout, the per-step
data recipe, the held-out split and the audit fallback."""
import numpy as np

from forkerrenderer_amd import nao as NA
from oracle import nao_ref as R


def test_param_layout_matches_oracle():
    assert NA.param_count() == R.NPARAM
    # 29,825 trained weights + 2,592 layer-1 weights of the zero padding channels
    assert R.NPARAM - 32 * 9 * 9 == 29825
    assert np.array_equal(NA.init_params(5), R.init_params(5))


def test_step_batch_is_a_function_of_seed_and_step():
    a = NA.step_batch(17, 3, 4000, 128, 16)
    b = NA.step_batch(17, 3, 4000, 128, 16)
    c = NA.step_batch(18, 3, 4000, 128, 16)
    for (fa, xa), (fb, xb) in zip(a, b):
        assert fa == fb and np.array_equal(xa, xb)
    assert [f for f, _ in a] != [f for f, _ in c]
    sizes = set()
    for step in range(40):
        for fs, xy in NA.step_batch(step, 3, 4000, 128, 16):
            W, H = NA.frame_size(fs)
            sizes.add((W, H))
            lW, lH = W // NA.SCALE, H // NA.SCALE
            assert xy.shape == (4, 2)
            assert (xy[:, 0] >= 0).all() and (xy[:, 0] <= lW - 128).all()
            assert (xy[:, 1] >= 0).all() and (xy[:, 1] <= lH - 128).all()
    assert sizes == {(1280, 800), (1920, 1080)}


def test_train_and_eval_frames_are_disjoint():
    tr = set(NA.frame_seeds("train", 4000))
    ev = set(NA.frame_seeds("eval", 200))
    assert len(tr) == 4000 and len(ev) == 200
    assert not tr & ev
    assert not {s // 1000 for s in tr} & {s // 1000 for s in ev}   # held-out layouts


def test_audit_fallback_replaces_only_bad_audited_tiles():
    H, W = 128, 256
    plain = np.full((H, W), 0.5, np.float32)
    pred = plain.copy()
    cover = np.ones((H, W), bool)
    pred[0:64, 64:128] += 0.2      # tile 1: bad
    pred[64:128, 0:64] += 0.01     # tile 4: within the threshold
    out, audited, failed = NA.audit_fallback(pred, plain, cover, 1, tile=64, every=2)
    # 2 x 4 tiles; frame 1 audits tiles 1, 3, 5, 7 (t % 2 == 1)
    assert audited == 4 and failed == 1
    assert np.array_equal(out[0:64, 64:128], plain[0:64, 64:128])
    out, audited, failed = NA.audit_fallback(pred, plain, cover, 0, tile=64, every=1)
    assert audited == 8 and failed == 1
    assert np.array_equal(out[64:128, 0:64], pred[64:128, 0:64])
    cover[:] = False
    _, audited, _ = NA.audit_fallback(pred, plain, cover, 0, tile=64, every=1)
    assert audited == 0
