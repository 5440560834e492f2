"""The package declares the reference's two console scripts (reference pyproject.toml:23-25),
and each resolves to a callable of the package."""

import importlib
from pathlib import Path

import tomli

ROOT = Path(__file__).resolve().parents[1]


def project():
    return tomli.loads((ROOT / "pyproject.toml").read_text())


def test_console_scripts_resolve():
    scripts = project()["project"]["scripts"]
    assert set(scripts) == {"qpop-diffraction", "qpop-diffraction-rsm"}
    for name, target in scripts.items():
        module, func = target.split(":")
        obj = getattr(importlib.import_module(module), func)
        assert callable(obj), name


def test_package_dir_is_the_package():
    cfg = project()["tool"]["setuptools"]
    src = ROOT / cfg["package-dir"]["qpop_diffraction"]
    assert (src / "__init__.py").is_file()
    assert cfg["packages"] == ["qpop_diffraction"]
    assert "lib/libqpd.so" in project()["tool"]["setuptools"]["package-data"]["qpop_diffraction"]


def test_rsm_command_parses_help(capsys):
    from qpop_diffraction.rsm import main
    try:
        main(["--help"])
    except SystemExit as e:
        assert e.code == 0
    assert "usage" in capsys.readouterr().out.lower()
