"""Mutation check of the oracle's pins (CPU only, run by hand; not collected
by pytest): plants one-symbol mistakes, one at a time, in a scratch copy of
oracle/hair_oracle.cpp, rebuilds it there and runs the oracle tests, which
must fail for each. The copy lives in a temporary directory; the repository
is not modified.

    python tests/oracle_mutations.py [mutation name ...]

"filter promote rule" is an equivalent mutant (a pixel with a centre hit
always weighs its own tap at 1, so the promotion test never decides there)
and is expected to survive.
"""
import os
import shutil
import subprocess
import sys
import tempfile

SRC = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DST = os.path.join(tempfile.mkdtemp(prefix="oracle_mut_"), "repo")
MUTS = [
 ("a_back exponent", "ab3 * af2 / (rest * rest * rest)", "ab3 * af2 / (rest * rest)"),
 ("delta_b sign", "alpha_b * (splat3(T(1)) - T(2) * ab2", "alpha_b * (splat3(T(1)) + T(2) * ab2"),
 ("delta_b dropped factor", "T(4) * af2 * ab2", "T(4) * af2"),
 ("sigma_b 3->2", "T(2) * beta_f2 + T(3) * beta_b2", "T(2) * beta_f2 + T(2) * beta_b2"),
 ("sigma_b af2->a_f", "(splat3(T(1)) + T(0.7f) * af2)", "(splat3(T(1)) + T(0.7f) * a_f)"),
 ("alpha_b from fwd", "v3<T> alpha_b = lobe_mean(back, a_b, shape.alpha);", "v3<T> alpha_b = lobe_mean(fwd, a_b, shape.alpha);"),
 ("global lobe swap", "gaussian3(vsqrt(splat3(beta2.y) + sigma_f2), x - splat3(shape.alpha.y)) * ds.fwd.tt", "gaussian3(vsqrt(splat3(beta2.z) + sigma_f2), x - splat3(shape.alpha.y)) * ds.fwd.tt"),
 ("global drop transmit", "out.global = (T(1) - direct) * transmit * T(0.7f)", "out.global = (T(1) - direct) * T(0.7f)"),
 ("global back sign", "gaussian3(vsqrt(ds.sigma_b2 + sigma_f2), x - ds.delta_b)", "gaussian3(vsqrt(ds.sigma_b2 + sigma_f2), x + ds.delta_b)"),
 ("global sigma_f2", "v3<T> sigma_f2 = depth * ds.beta_f2;", "v3<T> sigma_f2 = ds.beta_f2;"),
 ("local back 2->1", "v3<T> back = T(2) * ds.a_back", "v3<T> back = T(1) * ds.a_back"),
 ("dom slack", "thickness * T(0.05f), T(0));", "thickness * T(0.5f), T(0));"),
 ("dom eq8 sign", "xyz(p.light_forward)) -\n                 (-std::log", "xyz(p.light_forward)) +\n                 (-std::log"),
 ("dom bilinear tx/ty", "return mixs(mixs(v00, v10, tx), mixs(v01, v11, tx), ty);", "return mixs(mixs(v00, v10, ty), mixs(v01, v11, tx), ty);"),
 ("dom sigma", "return std::exp(-optical_depth(pix, rel / thickness) * p.filter_params.w);", "return std::exp(-optical_depth(pix, rel / thickness) * p.filter_params.z);"),
 ("spot y flag", "if (L.flags & HM_LIGHT_FLAG_SPOT_Y_NEGATIVE) spot_dir.y = -spot_dir.y;", ""),
 ("spot squared", "    return att * att;\n}\n\nstruct Lights", "    return att;\n}\n\nstruct Lights"),
 ("attenuation window", "T smooth = clampf(T(1) - factor * factor", "T smooth = clampf(T(1) - factor"),
 ("reconstruct y flip", "T ndc_y = (T(1) - uy) * T(2) - T(1);", "T ndc_y = uy * T(2) - T(1);"),
 ("roughness uvw", "in.roughness = fmax_(p.appearance.x * (T(0.65f) + T(0.7f) * uvw.z)", "in.roughness = fmax_(p.appearance.x * (T(0.65f) + T(0.7f) * uvw.y)"),
 ("albedo factor", "in.albedo = xyz(p.albedo) * (T(0.65f) + T(0.35f) * uvw.z);", "in.albedo = xyz(p.albedo) * (T(0.65f) + T(0.3f) * uvw.z);"),
 ("filter perp sigma", "d_perp * d_perp / (sigma_perp * sigma_perp)", "d_perp * d_perp / (sigma_par * sigma_perp)"),
 ("filter chroma", "std::exp(-chroma * chroma / (sigma_c * sigma_c))", "std::exp(-chroma / (sigma_c * sigma_c))"),
 ("filter depth reject", "depth_reject = T(1.45e-3f)", "depth_reject = T(1.45e-2f)"),
 ("filter promote rule", "if (promote && weight <= T(1e-5f))", "if (weight <= T(1e-5f))"),
 ("filter closer", "(center_depth - cons_depth) >= closer_eps", "(cons_depth - center_depth) >= closer_eps"),
 ("screen tangent swap", "return {d.x / l, d.y / l};", "return {d.y / l, d.x / l};"),
 ("every caster gets DOM", "if (dom_left && casts) {", "if (casts) {"),
 ("spot mask on points", "if (pass == 1) c = c * spot_mask<T>(L, world);", "if (pass == 0) c = c * spot_mask<T>(L, world);"),
 ("ortho view sign", "if (in.ortho) in.view = safe_normalize(v3<T>{p.clip[2], p.clip[6], p.clip[10]});", "if (in.ortho) in.view = safe_normalize(v3<T>{p.clip[1], p.clip[5], p.clip[9]});"),
 ("cube +x sc sign", "sc = d.x >= T(0) ? -d.z : d.z;", "sc = d.x >= T(0) ? d.z : d.z;"),
 ("cube +y tc sign", "tc = d.y >= T(0) ? d.z : -d.z;", "tc = d.y >= T(0) ? -d.z : -d.z;"),
 ("cube texel offset", "T u = c.s * T((float)size) - T(0.5f)", "T u = c.s * T((float)size)"),
 ("cube bilinear weights", "return mix3(mix3(tap[0], tap[1], wx), mix3(tap[2], tap[3], wx), wy);", "return mix3(mix3(tap[0], tap[1], wy), mix3(tap[2], tap[3], wx), wy);"),
 ("cube level blend", "return mix3(a, cube_bilinear(lv[1], sz[1], c), f);", "return mix3(a, cube_bilinear(lv[1], sz[1], c), T(1) - f);"),
 ("cubemap no z flip", "    s.z = -s.z;  // cube maps are left-handed\n", ""),
 ("quat rotate sign", "return d + T(q[3]) * t + cross3(u, t);", "return d - T(q[3]) * t + cross3(u, t);"),
 ("probe tilt sign", "T theta_i = tilt - theta_o;", "T theta_i = tilt + theta_o;"),
 ("probe trt half angle", "T f_trt = fresnel(cos_o * T(0.5f));", "T f_trt = fresnel(cos_o);"),
 ("probe trt absorption", "v3<T> e = sigma_a * (T(-3.2f) / cos_o);", "v3<T> e = sigma_a * (T(-2.0f) / cos_o);"),
 ("probe tint shadow", "scatter_tint(albedo, T(0.5f))", "scatter_tint(albedo, T(0.0f))"),
 ("probe level", "roughness * T((float)(env->specular_mips - 1u))", "roughness * T((float)env->specular_mips)"),
 ("probe no ao", "p.appearance.y) * in.ao;", "p.appearance.y);"),
 ("rect normal order", "safe_normalize(cross3(v3<T>{T(L.up[0]), T(L.up[1]), T(L.up[2])},\n                                             v3<T>{T(L.right[0]), T(L.right[1]), T(L.right[2])}))", "safe_normalize(cross3(v3<T>{T(L.right[0]), T(L.right[1]), T(L.right[2])},\n                                             v3<T>{T(L.up[0]), T(L.up[1]), T(L.up[2])}))"),
 ("rect solid angle", "T solid = T(L.width) * T(L.height) * facing / dist2;", "T solid = T(L.width) * T(L.height) * facing;"),
 ("rect range", "T range2 = fmax_(T(L.range) * T(L.range), T(1e-4f));", "T range2 = fmax_(T(L.range), T(1e-4f));"),
]
TESTS = ["tests/test_oracle.py", "tests/test_oracle_pins.py", "tests/test_branch_cases.py", "tests/test_env_pins.py"]
shutil.copytree(SRC, DST, ignore=shutil.ignore_patterns('.git', '_work', 'gpurun_out', '__pycache__'))
orig = open(f'{SRC}/oracle/hair_oracle.cpp').read()
missed = []
only = sys.argv[1:]
for name, a, b in MUTS:
    if only and name not in only: continue
    n = orig.count(a)
    if n != 1:
        print(f"!! {name}: pattern found {n} times"); missed.append(name); continue
    open(f'{DST}/oracle/hair_oracle.cpp', 'w').write(orig.replace(a, b))
    subprocess.run(['make', '-s', '-C', f'{DST}/oracle', 'clean', 'all'], check=True, capture_output=True)
    r = subprocess.run([sys.executable, '-m', 'pytest', '-x', '-q', '-p', 'no:cacheprovider', *TESTS], cwd=DST, capture_output=True, text=True)
    line = [l for l in r.stdout.splitlines() if l.startswith('FAILED')][:1]
    caught = r.returncode != 0 and bool(line)  # an assertion failed (not a collection error)
    print(f"{'caught' if caught else 'MISSED'}: {name} {line[0][:110] if line else ''}", flush=True)
    if not caught: missed.append(name)
print("missed:", missed)
shutil.rmtree(os.path.dirname(DST), ignore_errors=True)
sys.exit(0 if missed in ([], ["filter promote rule"]) else 1)
