"""Writer of synthetic MBK .vps files for the loader tests: the layout LoadVPS reads
(vps_loader.h:87-247 over the job-file model of JobParam.h:246-384), with comments in every
form the reader strips.  The arrays written are returned, so a test checks what the engine
reads against what was written, not against a second reading of the file."""
import numpy as np


def radial_grid(n=240, r_min=1.0e-4, r_max=12.0):
    xi = np.linspace(np.log(r_min), np.log(r_max), n)
    return xi, np.exp(xi)


def synth_pp(n=240, channels=((0, 1.0, 0.9, 0.6, 1.6), (1, -0.5, -0.45, 0.55, 1.8)), pcc=True,
             seed=5, extra_nandl=(), scale=1.0):
    """channels: (l, E_up, E_down, sigma, cutoff) per projector; the two j-components are
    Gaussians of slightly different widths, times scale.  Returns the dict of arrays the file
    will hold."""
    rng = np.random.default_rng(seed)
    xi, r = radial_grid(n)
    vloc = -4.0 / np.sqrt(r * r + 0.64)
    comps, ls, eu, ed, nandl = [], [], [], [], []
    for k, (l, e_up, e_dn, sigma, cutoff) in enumerate(channels):
        amp = scale * (1.0 + 0.1 * k)
        up = amp * r ** l * np.exp(-r * r / (2 * sigma * sigma))
        dn = amp * r ** l * np.exp(-r * r / (2 * (1.05 * sigma) ** 2)) * (1.0 + 0.02 * rng.standard_normal())
        comps += [up, dn]
        ls.append(l)
        eu.append(e_up)
        ed.append(e_dn)
        nandl.append((k + 1, 2 + l, l, cutoff))
    nandl += list(extra_nandl)
    d = dict(xi=xi, r=r, vloc=vloc, comps=comps, l=ls, e_up=eu, e_dn=ed, nandl=nandl, valence=4.0,
             local_cutoff=2.2, pcc=(0.3 * np.exp(-r * r) if pcc else None))
    return d


def write_vps(path, d, vps_type="MBK", close_potentials=True, energies_count=None, pcc_lines=None,
              duplicate_keys=False):
    f = "{:.17e}".format
    out = []
    out.append("/* synthetic MBK pseudopotential for the loader tests")
    out.append("   lines inside this comment are skipped: vps.type NOT_MBK")
    out.append("*/")
    out.append("# header comment")
    out.append(f"vps.type   {vps_type}    # the format")
    if duplicate_keys:
        out.append("vps.type   OTHER")
    out.append(f"valence.electron {d['valence']}  <VPS> trailing comment word")
    out.append(f"local.cutoff\t{d['local_cutoff']} * star comment")
    if duplicate_keys:
        out.append("local.cutoff 99.0")
    out.append("<pseudo.NandL   * index n l cutoff")
    for (i, n, l, c) in d["nandl"]:
        out.append(f"   {i}  {n}  {l}  {c}")
    out.append("pseudo.NandL>")
    out.append("")
    npj = len(d["l"]) if energies_count is None else energies_count
    out.append("<project.energies")
    out.append(f"  {npj}")
    for l, a, b in zip(d["l"], d["e_up"], d["e_dn"]):
        out.append(f"  {l}  {f(a)}  {f(b)}   # l E_up E_down")
    out.append("project.energies>")
    out.append("<Pseudo.Potentials")
    for i in range(len(d["r"])):
        vals = [d["xi"][i], d["r"][i], d["vloc"][i]] + [c[i] for c in d["comps"]]
        out.append("  " + "  ".join(f(v) for v in vals))
    if close_potentials:
        out.append("Pseudo.Potentials>")
    if d["pcc"] is not None and close_potentials:
        out.append("<density.PCC")
        n = len(d["r"]) if pcc_lines is None else pcc_lines
        for i in range(n):
            out.append(f"  {f(d['xi'][i])}  {f(d['r'][i])}  {f(d['pcc'][i])}")
        out.append("density.PCC>")
    with open(path, "w") as fh:
        fh.write("\n".join(out) + "\n")
