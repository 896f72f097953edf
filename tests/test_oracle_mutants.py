"""Each pin of tests/oracle_pins.py is shown to catch the mistakes it is there
for: one-token mistakes are planted in copies of the oracle's sources (on the
CPU, in a temporary directory), each copy is compiled, and the pins must fail
on every mutant while passing on an unmodified copy built the same way."""
import concurrent.futures as cf
import json
import shutil
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
ORACLE = ROOT / "oracle"
TESTS = ROOT / "tests"

# (file, text in the oracle, planted replacement, pin expected to catch it)
MUTANTS = [
    # fresnel_conductor (bxdf.rs:190-217)
    ("integrator.cpp", "RGB t0 = eta2 - etak2 - s2;", "RGB t0 = eta2 - etak2 + s2;", "pin_fresnel_conductor_complex_index"),
    ("integrator.cpp", "sqrt_rgb(t0 * t0 + eta2 * etak2 * 4.0f)", "sqrt_rgb(t0 * t0 + eta2 * etak2 * 2.0f)",
     "pin_fresnel_conductor_normal_incidence"),
    ("integrator.cpp", "RGB rp = rs * (t3 - t4) / (t3 + t4);", "RGB rp = rs * (t3 + t4) / (t3 - t4);",
     "pin_fresnel_conductor_complex_index"),
    ("integrator.cpp", "RGB eta = eta_t / eta_i;\n    RGB eta_k = k / eta_i;",
     "RGB eta = k / eta_i;\n    RGB eta_k = eta_t / eta_i;", "pin_fresnel_conductor_normal_incidence"),
    ("integrator.cpp", "return (rp + rs) * 0.5f;", "return rs;", "pin_fresnel_conductor_complex_index"),
    ("integrator.cpp", "RGB t2 = c * a * 2.0f;", "RGB t2 = c * a;", "pin_fresnel_conductor_normal_incidence"),
    # Trowbridge-Reitz D, Lambda, G1, pdf (microfacet.rs:17-29, 116-142)
    ("integrator.cpp", "Float a2t2 = (alpha * att) * (alpha * att);", "Float a2t2 = alpha * att * att;",
     "pin_tr_lambda_closed_form"),
    ("integrator.cpp", "cos4 * (1.0f + e) * (1.0f + e));", "cos4 * (1.0f + e));", "pin_tr_visible_normals_normalised"),
    ("integrator.cpp", "return 1.0f / (1.0f + tr_lambda(d, w));", "return 1.0f / (1.0f + 2.0f * tr_lambda(d, w));",
     "pin_tr_visible_normals_normalised"),
    ("integrator.cpp", "return 1.0f / (1.0f / tr_lambda(d, w));", "return 1.0f / (1.0f + tr_lambda(d, w));",
     "pin_tr_ref_pdf_ratio"),
    ("integrator.cpp", "abs_dot(wo, wh) / abs_cos_theta(wo); }", "abs_dot(wo, wh); }", "pin_tr_visible_normals_normalised"),
    # sample11 and sample (microfacet.rs:55-153)
    ("integrator.cpp", "Float r = std::sqrt(u1 / (1.0f - u1));", "Float r = std::sqrt(u1);",
     "pin_sample11_normal_incidence"),
    ("integrator.cpp", "else slope_x = (mode == MODE_REF) ? x1 : x2;", "else slope_x = (mode == MODE_REF) ? x1 : x1;",
     "pin_tr_sample_follows_visible_normals"),
    ("integrator.cpp", "else slope_x = (mode == MODE_REF) ? x1 : x2;", "else slope_x = (mode == MODE_REF) ? x2 : x2;",
     "pin_sample11_reference_keeps_the_smaller_root"),
    # (swapping the test a < 0 for a > 0 only relabels which u1 gives which root: the sampled
    # distribution is the same, so no distribution test can see it; the same-stream GPU-oracle
    # renders pin that choice instead)
    ("integrator.cpp", "if (a < 0.0f || x2 > 1.0f / tan_t) slope_x = x1;", "if (a < 0.0f || x2 < 1.0f / tan_t) slope_x = x1;",
     "pin_tr_sample_follows_visible_normals"),
    ("integrator.cpp", "Float tmp = cos_phi(ws) * sx - sin_phi(ws) * sy;", "Float tmp = cos_phi(ws) * sx + sin_phi(ws) * sy;",
     "pin_tr_sample_follows_visible_normals"),
    ("integrator.cpp", "V3 ws = normalize(V3(ax * wi.x, ay * wi.y, wi.z));", "V3 ws = normalize(V3(wi.x, wi.y, wi.z));",
     "pin_tr_sample_follows_visible_normals"),
    ("integrator.cpp", "return flip ? -wh : wh;", "return wh;", "pin_tr_sample_flip"),
    # offset_ray_origin (point3.rs:31-49), spawn_ray / spawn_ray_to_it (interaction.rs:115-143)
    ("rt.h", "if (dot(w, n) < 0.0f) offset = -offset;", "if (dot(w, n) > 0.0f) offset = -offset;", "pin_spawn_ray"),
    ("rt.h", "if (offset[i] > 0.0f) po.at(i) = next_float_up(po[i]);",
     "if (offset[i] > 0.0f) po.at(i) = next_float_down(po[i]);", "pin_spawn_ray"),
    ("rt.h", "V3 offset = dot(vabs(n), perr) * n;", "V3 offset = dot(n, perr) * n;", "pin_spawn_ray"),
    ("rt.h", "else if (offset[i] < 0.0f) po.at(i) = next_float_down(po[i]);",
     "else if (offset[i] < 0.0f) po.at(i) = po[i];", "pin_spawn_ray"),
    ("integrator.cpp", "mode == MODE_REF ? 1.0f - 0.0001f : INFINITY", "mode == MODE_REF ? 1.0f - 0.001f : INFINITY",
     "pin_spawn_ray"),
    ("integrator.cpp", "V3 t = offset_ray_origin(p1, e1, n1, o - p1);", "V3 t = offset_ray_origin(p1, e1, n1, p1 - o);",
     "pin_spawn_ray_to"),
    ("integrator.cpp", "V3 o = offset_ray_origin(p0, e0, n0, p1 - p0);", "V3 o = offset_ray_origin(p0, e0, n0, p0 - p1);",
     "pin_spawn_ray_to"),
    ("integrator.cpp", "return Ray{o, t - o, 1.0f - 0.0001f, time};", "return Ray{o, t - o, 1.0f, time};",
     "pin_spawn_ray_to"),
]


def _build(k: int, tmp: Path, mutant) -> tuple[int, str]:
    d = tmp / f"m{k}"
    shutil.copytree(ORACLE, d, ignore=shutil.ignore_patterns("*.so", "_ref", "__pycache__", "*.py"))
    if mutant is not None:
        f, old, new, _ = mutant
        src = (d / f).read_text()
        assert src.count(old) == 1, (f, old)
        (d / f).write_text(src.replace(old, new))
    out = d / f"liboracle_m{k}.so"
    cmd = ["g++", "-O1", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-pthread", "-shared",
           "-o", str(out), *map(str, sorted(d.glob("*.cpp")))]
    subprocess.run(cmd, check=True, capture_output=True)
    r = subprocess.run([sys.executable, str(TESTS / "oracle_pins.py"), str(out)], capture_output=True, text=True,
                       cwd=str(TESTS), timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    return k, r.stdout.strip().splitlines()[-1]


@pytest.fixture(scope="module")
def results(tmp_path_factory):
    tmp = tmp_path_factory.mktemp("mutants")
    jobs = [None] + MUTANTS
    with cf.ThreadPoolExecutor(8) as ex:
        out = dict(ex.map(lambda kv: _build(kv[0], tmp, kv[1]), enumerate(jobs)))
    return {k: json.loads(v) for k, v in out.items()}


def test_unmodified_copy_passes_every_pin(results):
    assert results[0] == []


@pytest.mark.parametrize("k", range(1, len(MUTANTS) + 1), ids=[f"{m[0]}:{m[1][:40]}" for m in MUTANTS])
def test_planted_mistake_is_caught(results, k):
    f, old, new, pin = MUTANTS[k - 1]
    assert pin in results[k], (old, new, results[k])
