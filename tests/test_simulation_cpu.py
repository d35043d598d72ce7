"""CPU checks of the legacy-simulation host logic: configuration checks and
the builder's refusals of scheme choices that are not built (no device
work happens before them)."""

import pytest

from twophaseflow_amd import NumericsConfig, SimulationBuilder, SimulationConfig


@pytest.mark.parametrize("kw", [dict(),  # the reference's default fccd_flux advection
                                dict(advection_scheme="dissipative_ccd", convection_scheme="ccd"),
                                dict(advection_scheme="dissipative_ccd",
                                     surface_tension_model="gfm"),
                                dict(advection_scheme="dissipative_ccd", reinit_method="eikonal"),
                                dict(advection_scheme="dissipative_ccd",
                                     extension_method="upwind")])
def test_builder_rejects_unbuilt_schemes(kw):
    with pytest.raises(NotImplementedError):
        SimulationBuilder(SimulationConfig(numerics=NumericsConfig(**kw))).build()


def test_numerics_config_checks():
    with pytest.raises(AssertionError):
        NumericsConfig(advection_scheme="weno5")
    with pytest.raises(AssertionError):
        NumericsConfig(uccd6_sigma=0.0)
    assert NumericsConfig(advection_scheme="fccd").advection_scheme == "fccd_flux"
