"""Transcribe the reference's committed numeric outputs into a JSON fixture.

Sources (Arth-Singh/geon):
  docs/energy_conditions.txt:1-49  output of src/bin/energy_conditions.rs,
                                   values printed with Rust's {:+.4e}
  README.md:42-50                  output of src/bin/validate_shadow.rs and
                                   the Minkowski null-residual claim

Run once in a checkout that has the reference at /root/reference:
    python tests/golden/make_reference_anchors.py
It writes tests/golden/reference_anchors.json (data only: the inputs of each
line and the printed value, as text and as float).
"""
import json
import os
import re

REF = "/root/reference"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_anchors.json")


def parse_energy_conditions(path):
    text = open(path, encoding="utf-8").read().splitlines()
    section = None
    rows = []
    num = r"([+-]?\d+\.\d+e[+-]?\d+)"
    for ln, line in enumerate(text, start=1):
        if line.startswith("== "):
            section = line.split("(")[0].strip("= ").strip()
            continue
        m = re.match(r"\s*r =\s+(\d+)\s+(NEC contraction|Ricci scalar)\s+" + num, line)
        if m:
            rows.append(dict(line=ln, section=section, kind="schwarzschild_" +
                             ("nec" if m.group(2).startswith("NEC") else "ricci_scalar"),
                             coord=float(m.group(1)), text=m.group(3), value=float(m.group(3))))
            continue
        m = re.match(r"\s*(T_μν k\^μ k\^ν|Ricci scalar)\s+" + num, line)
        if m and section and section.startswith("Minkowski"):
            rows.append(dict(line=ln, section=section, kind="minkowski_" +
                             ("nec" if m.group(1).startswith("T") else "ricci_scalar"),
                             coord=None, text=m.group(2), value=float(m.group(2))))
            continue
        m = re.match(r"\s*([xl]) =\s+([-\d.]+)\s+" + num, line)
        if m:
            kind = "alcubierre_nec" if m.group(1) == "x" else "morris_thorne_nec"
            rows.append(dict(line=ln, section=section, kind=kind, coord=float(m.group(2)),
                             text=m.group(3), value=float(m.group(3))))
            continue
        m = re.match(r"\s*minimum T_μν k\^μ k\^ν = " + num + r" at ([xl]) = ([-\d.]+)", line)
        if m:
            kind = "alcubierre_nec_min" if m.group(2) == "x" else "morris_thorne_nec_min"
            rows.append(dict(line=ln, section=section, kind=kind, coord=float(m.group(3)),
                             text=m.group(1), value=float(m.group(1))))
            continue
        m = re.match(r"\s*∫ T_kk dx .*≈ " + num, line)
        if m:
            rows.append(dict(line=ln, section=section, kind="alcubierre_nec_integral",
                             coord=None, text=m.group(1), value=float(m.group(1))))
            continue
        m = re.match(r"\s*T\[(\d)\]\[(\d)\] = " + num, line)
        if m:
            rows.append(dict(line=ln, section=section, kind="morris_thorne_T_diag_throat",
                             coord=int(m.group(1)), text=m.group(3), value=float(m.group(3))))
    return rows


def parse_readme(path):
    text = open(path, encoding="utf-8").read().splitlines()
    out = {}
    for ln, line in enumerate(text, start=1):
        m = re.match(r"measured: b_c =\s+([\d.]+), α_c = ([\d.]+)°", line.strip())
        if m:
            out["validate_shadow_measured"] = dict(line=ln, b_c_text=m.group(1),
                                                   alpha_c_deg_text=m.group(2))
        m = re.match(r"analytic: b_c = 3√3 M = ([\d.]+), α_c = ([\d.]+)°", line.strip())
        if m:
            out["validate_shadow_analytic"] = dict(line=ln, b_c_text=m.group(1),
                                                   alpha_c_deg_text=m.group(2))
        m = re.match(r"relative error: ([\d.]+e-?\d+)", line.strip())
        if m:
            out["validate_shadow_rel_err"] = dict(line=ln, text=m.group(1))
        m = re.search(r"\|g_μν p\^μ p\^ν\| ≈ ([\d.e-]+)", line)
        if m:
            out["minkowski_null_residual"] = dict(line=ln, text=m.group(1))
    return out


def main():
    rows = parse_energy_conditions(os.path.join(REF, "docs", "energy_conditions.txt"))
    readme = parse_readme(os.path.join(REF, "README.md"))
    fixture = dict(
        source="Arth-Singh/geon docs/energy_conditions.txt and README.md",
        energy_conditions=dict(
            inputs=dict(
                minkowski=dict(x=[0.0, 1.0, 1.0, 1.0], k=[1.0, 1.0, 0.0, 0.0]),
                schwarzschild=dict(mass=1.0, theta="pi/2", k="(1/sqrt(f), -sqrt(f), 0, 0)"),
                alcubierre=dict(bubble_radius=2.0, bubble_thickness=0.5, ship_velocity=2.0,
                                ship_x=0.0, k=[1.0, 1.0, 0.0, 0.0],
                                scan="x=-4+0.2n, n=0..40, printed every 4th"),
                morris_thorne=dict(throat_radius=1.0, k=[1.0, 1.0, 0.0, 0.0],
                                   scan="l=-4+0.2n, theta=pi/2, phi=0"),
            ),
            rows=rows),
        readme=readme,
        validate_shadow_inputs=dict(mass=1.0, r_obs=20.0, step=0.01, max_steps=200000,
                                    iterations=60, alpha_range="[0, pi/2-1e-3]"),
    )
    with open(OUT, "w", encoding="utf-8") as fh:
        json.dump(fixture, fh, indent=1, ensure_ascii=False)
    print(f"wrote {OUT}: {len(rows)} energy-condition rows, readme keys {sorted(readme)}")


if __name__ == "__main__":
    main()
