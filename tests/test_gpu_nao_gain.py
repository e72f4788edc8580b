"""What the trained network gains over the method without it: on held-out
seeded frames of the 1280x800 generator (layouts and camera poses the
training never saw), the shipped trained weights against the same network
untrained (its initialisation) and against SSAO without its blur (the
cheaper plain alternative), all measured against plain SSAO + blur on the
covered pixels.  The trained network must stay within the AO error bound
(MAE <= 0.02), be far better than untrained, and better than unblurred
SSAO."""
import pytest

from forkerrenderer_amd import nao as NA

pytestmark = pytest.mark.gpu

FRAMES = 12


@pytest.mark.timeout(600)
def test_trained_network_beats_untrained_and_unblurred_ssao(ctx, tmp_path):
    sset = NA.SceneSet(ctx, str(tmp_path / "sets"))
    ck = NA.load_checkpoint(NA.DEFAULT_WEIGHTS)
    trained = NA.NeuralAO(ctx, batch=1, crop=64, params=ck["state"][:NA.param_count()], scale=int(ck["scale"]))
    untrained = NA.NeuralAO(ctx, batch=1, crop=64, params=NA.init_params(0), scale=int(ck["scale"]))
    try:
        print(f"held-out frames: {FRAMES}, trained weights", flush=True)
        t = NA.evaluate(trained, ctx, sset, FRAMES, unblurred=True)
        print("held-out frames: untrained", flush=True)
        u = NA.evaluate(untrained, ctx, sset, FRAMES)
    finally:
        trained.close()
        untrained.close()
    print(f"MAE trained {t['mae']:.4f}, untrained {u['mae']:.4f}, unblurred SSAO {t['mae_unblurred_ssao']:.4f}; "
          f"fallback rate trained {t['fallback_rate']:.3f}, untrained {u['fallback_rate']:.3f}", flush=True)
    assert t["mae"] <= 0.02, t
    assert t["mae"] < 0.5 * u["mae"], (t, u)
    assert t["mae"] < t["mae_unblurred_ssao"], t
    assert t["fallback_rate"] < u["fallback_rate"] or u["fallback_rate"] == 0, (t, u)
