"""The reference's validation run (runValidation src/main.cpp:4220-4583)
through fire_validate: the same CSV columns and JSON report keys, target
envelope evaluation, the stability checks, and the BMP output."""
import csv
import json
import os
import struct
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(os.path.dirname(HERE), "native-firesim_amd", "fire_validate")

# validation-metrics.csv header of the reference (src/main.cpp:4254-4260)
COLUMNS = ("frame,timeSeconds,gridX,gridY,gridZ,pressureIterations,gpuSolveMs,gpuRenderMs,gpuVelocityMs,gpuReactionMs,"
           "gpuProjectionMs,gpuLightingMs,gpuRaymarchMs,gpuPackMs,heatSum,fuelSum,oxygenSum,sootSum,charSum,ashSum,"
           "pyrolysisSum,progressSum,turbulenceEnergySum,sootOpticalDepthSum,maxHeat,maxFuel,maxSoot,maxPyrolysis,"
           "maxProgress,maxTurbulenceEnergy,flameHeightMeters,meanOpticalDepth,meanSceneLight,meanSceneShadow,"
           "flameMassProxy,smokeMassProxy,flameSmokeOverlapProxy,meanVolumetricShadow,meanRoomIrradiance,"
           "heatReleaseProxy,divergenceBeforeL2,divergenceAfterL2,divergenceBeforeMax,divergenceAfterMax,"
           "divergenceReduction,invalidCells").split(",")


def test_validator_is_built():
    assert os.access(EXE, os.X_OK)


@pytest.mark.gpu
def test_validation_run_passes_and_writes_reference_artefacts(hip_lib, tmp_path):
    targets = tmp_path / "targets.csv"
    targets.write_text("# metric,min,max\nmetric,min,max\nworstDivergenceAfterL2,0.0,0.35\n"
                       "averageDivergenceReduction,0.02,1.0\nimageMaxLuma,0.12,1.0\naverageGpuRenderMs,0,110\n")
    out = tmp_path / "out"
    r = subprocess.run([EXE, f"--targets={targets}", f"--output-dir={out}"], capture_output=True, text=True, timeout=110)
    rep = json.loads((out / "validation-report.json").read_text())
    assert r.returncode == 0, (r.stderr, {k: rep[k] for k in ("maxHeat", "maxFlameHeightMeters", "imageBrightPixels")})
    assert rep["validationOk"] is True and rep["frames"] == 96 and rep["grid"] == [176, 208, 128]
    assert rep["pressureIterations"] == 40
    assert [t["passed"] for t in rep["targetEnvelopes"]] == [True] * 4
    rows = list(csv.reader(open(out / "validation-metrics.csv")))
    assert rows[0] == COLUMNS and len(rows) == 97
    assert all(int(r[-1]) == 0 for r in rows[1:])          # invalidCells
    assert [int(r[0]) for r in rows[1:]] == list(range(96))  # frame index from the in-call reset
    data = open(out / "validation-frame.bmp", "rb").read()
    assert struct.unpack_from("<HI", data, 0) == (0x4D42, 54 + 960 * 540 * 4)


@pytest.mark.gpu
def test_unmatched_target_fails_validation(hip_lib, tmp_path):
    targets = tmp_path / "targets.csv"
    targets.write_text("calibrationHrrShapeRmse,0,1\n")
    out = tmp_path / "out"
    r = subprocess.run([EXE, f"--targets={targets}", f"--output-dir={out}", "--validation-frames=8"], capture_output=True,
                       text=True, timeout=110)
    assert r.returncode == 4
    rep = json.loads((out / "validation-report.json").read_text())
    assert rep["validationOk"] is False
    assert rep["targetEnvelopes"][0]["matched"] is False
