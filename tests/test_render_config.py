"""The render-config importer (render_config_importer.cc:47-145) against
hand-written configurations: every key of both integrator types, the
directory form of the file name, the clamp modes, the errors for what the
reference leaves unset, and the mapping onto the engine's parameters."""
import json

import pytest

from renzoku_amd import _lib
from renzoku_amd.render_config import (RenderConfigError, config_file, load_render_config,
                                       parse_render_config)

VPL = {"integrator": {"type": "vpl", "max_bounce": 5, "direct_lighting": True,
                      "radiance_estimate": "lightslice_guided_path_trace",
                      "lightslice": {"clusters": 300, "max_clusters": 600, "slice_size": 256, "density_radius": 0.25,
                                     "neighbor_slices": 16},
                      "start_particles": 65536, "clamp_type": "radiance",
                      "clamp_threshold": {"radiance": 0.05, "incident_radiance": 0.5}, "clamp_relaxation": 0.8}}


def test_vpl_integrator_every_key():
    c = parse_render_config(VPL)
    assert c.integrator == "vpl" and c.max_bounce == 5 and c.direct_lighting is True
    assert c.radiance_estimate == "lightslice_guided_path_trace"
    assert (c.lightslice.clusters, c.lightslice.max_clusters, c.lightslice.slice_size,
            c.lightslice.density_radius, c.lightslice.neighbor_slices) == (300, 600, 256, 0.25, 16)
    assert c.start_particles == 65536
    assert c.clamp_mode == _lib.CLAMP_RADIANCE
    assert (c.clamp_threshold, c.incident_clamp_threshold, c.clamp_relaxation) == (0.05, 0.5, 0.8)


def test_path_integrator():
    c = parse_render_config({"integrator": {"type": "path", "max_bounce": 7, "samples": 64}})
    assert (c.integrator, c.max_bounce, c.samples) == ("path", 7, 64)
    assert c.path_params() == dict(spp=64, max_bounce=7, direct_lighting=1, guide_mode=0, seed=2013, tmin=1e-3)


def test_mapping_onto_engine_parameters():
    c = parse_render_config(VPL)
    kw = c.lightslice_params(512, 512, n_emitted=70000.0)
    assert kw["num_clusters"] == 300 and kw["max_slice_clusters"] == 600 and kw["kd_max_slice_size"] == 256
    assert kw["num_neighbor_slices"] == 16 and kw["kd_pivots"] == 0
    assert c.lightslice_params(512, 512, n_emitted=1.0, kd_pivots=1)["kd_pivots"] == 1
    assert kw["clamp_mode"] == 3 and kw["clamp"] == 0.05 and kw["incident_clamp"] == 0.5
    names = {f for f, _ in _lib.LightSliceParams._fields_}
    assert set(kw) <= names                       # every key is a field of rzk_lightslice_params
    pp = c.path_params(spp=16)
    assert pp["max_bounce"] == 5 and pp["direct_lighting"] == 1 and pp["guide_mode"] == 1 and pp["spp"] == 16
    assert set(pp) <= {f for f, _ in _lib.PtParams._fields_}


@pytest.mark.parametrize("clamp,keys,mode", [("none", {}, 0), ("form_factor", {"form_factor": 0.01,
                                                                               "incident_form_factor": 0.02}, 1),
                                             ("reflectivity", {"reflectivity": 1e-3,
                                                               "incident_reflectivity": 2e-3}, 2)])
def test_other_clamp_modes(clamp, keys, mode):
    d = json.loads(json.dumps(VPL))
    d["integrator"]["clamp_type"] = clamp
    d["integrator"]["clamp_threshold"] = keys
    c = parse_render_config(d)
    assert c.clamp_mode == mode
    if keys:
        assert c.clamp_threshold == keys[clamp] and c.incident_clamp_threshold == keys["incident_" + clamp]


def test_directory_form_and_files(tmp_path):
    d = tmp_path / "breakfast"
    d.mkdir()
    (d / "breakfast.json").write_text(json.dumps(VPL))
    assert config_file(str(d)) == str(d / "breakfast.json")
    assert config_file(str(d) + "/") == str(d / "breakfast.json")
    assert config_file(str(d / "other.json")) == str(d / "other.json")
    assert load_render_config(str(d)).lightslice.clusters == 300
    assert load_render_config(str(d / "breakfast.json")).max_bounce == 5
    (tmp_path / "bad.json").write_text("{ not json")
    with pytest.raises(RenderConfigError):
        load_render_config(str(tmp_path / "bad.json"))


@pytest.mark.parametrize("edit", [
    lambda d: d.pop("integrator"),
    lambda d: d["integrator"].update(type="photon"),
    lambda d: d["integrator"].update(radiance_estimate="lightcuts"),
    lambda d: d["integrator"].update(clamp_type="soft"),
    lambda d: d["integrator"].pop("max_bounce"),
    lambda d: d["integrator"].update(max_bounce=2.5),
    lambda d: d["integrator"].update(direct_lighting=1),
    lambda d: d["integrator"]["lightslice"].pop("neighbor_slices"),
    lambda d: d["integrator"]["clamp_threshold"].pop("incident_radiance"),
])
def test_what_the_reference_leaves_unset_is_an_error(edit):
    d = json.loads(json.dumps(VPL))
    edit(d)
    with pytest.raises(RenderConfigError):
        parse_render_config(d)
