"""CPU checks of the autotuner's bookkeeping: candidate forms per problem shape,
the JSON cache round trip, and the switch that turns tuning off."""

import importlib

import numpy as np
import pytest


@pytest.fixture
def tuner(tmp_path, monkeypatch):
    monkeypatch.setenv('DQH_AUTOTUNE_CACHE', str(tmp_path / 'tune.json'))
    from dynamiqs_amd import autotune

    importlib.reload(autotune)
    yield autotune
    importlib.reload(autotune)


def test_dia_candidates_follow_offsets(tuner):
    from dynamiqs_amd.integrators import dia

    # 32 x 32 two-mode operators: offsets up to 32 admit only the widest halo
    assert tuner.dia_candidates(1024, 32) == [dia.SWEEP_GLOBAL, dia.SWEEP_TILED_H32]
    assert tuner.dia_candidates(192, 12) == [dia.SWEEP_GLOBAL, dia.SWEEP_TILED_H16, dia.SWEEP_TILED_H32]
    assert tuner.dia_candidates(256, 2) == [dia.SWEEP_GLOBAL, dia.SWEEP_TILED_H8, dia.SWEEP_TILED_H16, dia.SWEEP_TILED_H32]
    assert tuner.dia_candidates(256, 36) == [dia.SWEEP_GLOBAL]


def test_cache_round_trip(tuner, tmp_path, monkeypatch):
    tuner.record('dia|dev|n=1024|nops=3|ndiag=5|maxoff=32', {'variant': 5, 'us': {'1': 300.0, '5': 120.0}})
    assert (tmp_path / 'tune.json').exists()
    import dynamiqs_amd.autotune as fresh

    importlib.reload(fresh)
    assert fresh.lookup('dia|dev|n=1024|nops=3|ndiag=5|maxoff=32')['variant'] == 5


def test_disabled_tuner_keeps_the_fixed_rule(tuner, monkeypatch):
    from dynamiqs_amd.integrators import dia

    monkeypatch.setenv('DQH_AUTOTUNE', '0')
    H = np.diag(np.arange(256.0)).astype(complex)
    assert tuner.dia_sweep(H, [], 4) == dia.SWEEP_AUTO
    assert tuner.dense_global(64, 2, 256) is False
