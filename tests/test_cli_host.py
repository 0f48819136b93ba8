"""The command-line driver's input handling: the reference's own singlet deck
(CSCALE-WORK/Input.Inp.singlet, copied as a fixture) parsed as a Fortran
namelist read does (last assignment wins), the input.F90 defaults, and the
get_args overrides (input.F90:264-362)."""
from doublyexcited_amd.cli import DEFAULTS, apply_args, load_config, parse_namelist


def test_reference_deck(golden_dir):
    cfg = parse_namelist((golden_dir / "Input.Inp.singlet").read_text())
    assert cfg["WRITEVECS"] is True and cfg["THETA"] == 0.2
    assert cfg["AUTODELTA"] is False and cfg["DELTA"] == 2.0
    assert cfg["LANTOL"] == 1e-2 and cfg["LANBLOCKSIZE"] == 3 and cfg["KDIM"] == 501
    assert cfg["GM_USEPETSCFLAG"] == 1 and cfg["GM_MAX_INNER_ITER"] == 20 and cfg["GM_TOL"] == 1e-4
    assert cfg["NUMORBS"] == 14 and cfg["SINGLET"] is True
    assert cfg["TAKEREALPART"] is True and cfg["GPLMR_SHIFT"] == complex(-0.76, 0.0)
    assert cfg["GPLMR_TOL"] == 1e-7 and cfg["GPLMR_NEV"] == 1 and cfg["OUTMODULUS"] == 10
    assert cfg["HERMOPT"] is False and cfg["CSHIFTOPT"] is False


def test_defaults_and_arguments(golden_dir, tmp_path):
    cfg = load_config([f"Inp={golden_dir / 'Input.Inp.singlet'}", "N=5", "Iter=5", "Notakereal", "Sing"])
    assert cfg["N"] == 5 and cfg["GM_MAX_INNER_ITER"] == 5 and cfg["TAKEREALPART"] is False
    assert cfg["DELTA"] == 2.0  # AUTODELTA off in the deck
    assert cfg["GPLMR_NS"] == DEFAULTS["GPLMR_NS"] == 5
    c2 = apply_args(cfg, ["R=10", "Theta=0.3", "Trip", "Shift=(-0.5,0.01)", "Nev=2", "Read"])
    assert c2["AUTODELTA"] is True and c2["RADIUS"] == 10.0 and c2["THETA"] == 0.3
    assert c2["SINGLET"] is False and c2["GPLMR_SHIFT"] == complex(-0.5, 0.01)
    assert c2["GPLMR_NEV"] == 2 and c2["READVECS"] is True and c2["NUMSKIP"] == 0
    # no deck: input.F90 defaults, AUTODELTA gives DELTA = RADIUS / N
    c3 = load_config([f"Inp={tmp_path / 'missing'}", "N=7"])
    assert c3["DELTA"] == 7.0 / 7
