"""Renzoku's render configuration file (RenderConfigImporter,
render_config_importer.cc:15-145) read into the engine's parameters.

The file is JSON with one top-level "integrator" object (load_integrator,
render_config_importer.cc:74-145):

* "type": "path" (PathTracing: "max_bounce", "samples") or "vpl"
  (VirtualPointLight: "max_bounce", "direct_lighting", "radiance_estimate",
  "lightslice" {"clusters", "max_clusters", "slice_size", "density_radius",
  "neighbor_slices"}, "start_particles", "clamp_type", "clamp_threshold",
  "clamp_relaxation").
* A path without ".json" names a configuration directory: <dir>/<name>/<name>.json
  (RenderConfigImporter::load, render_config_importer.cc:47-72).

The reference reads only the "lightslice_guided_path_trace" estimate and the
"radiance" clamp and leaves any other value unset; here those two read as in the
reference, the engine's other clamp modes ("none", "form_factor",
"reflectivity", thresholds under the keys "form_factor" / "reflectivity" and
"incident_form_factor" / "incident_reflectivity") are accepted as well, and any
other estimate, clamp or integrator type is an error rather than an unset value.
"density_radius" and "clamp_relaxation" are kept but nothing on the traced path
reads them.
"""
from __future__ import annotations

import json
import os
from dataclasses import dataclass, field

CLAMP_MODES = {"none": 0, "form_factor": 1, "reflectivity": 2, "radiance": 3}
ESTIMATES = ("lightslice_guided_path_trace",)


class RenderConfigError(ValueError):
    pass


@dataclass
class LightSliceConfig:
    clusters: int = 0               # set_num_clusters
    max_clusters: int = 0           # set_max_clusters_per_slice
    slice_size: int = 0             # set_max_slice_size (kd-tree slices of at most this many gather points)
    density_radius: float = 0.0     # set_density_radius
    neighbor_slices: int = 0        # set_num_neighbor_slices


@dataclass
class RenderConfig:
    integrator: str                 # "path" | "vpl"
    max_bounce: int
    samples: int = 0                # path: set_num_samples
    direct_lighting: bool = False   # vpl: set_direct_lighting
    radiance_estimate: str = ""
    lightslice: LightSliceConfig = field(default_factory=LightSliceConfig)
    start_particles: int = 0        # vpl: set_start_particles (light paths traced for VPLs)
    clamp_type: str = "form_factor"
    clamp_threshold: float = 0.0
    incident_clamp_threshold: float = 0.0
    clamp_relaxation: float = 0.0

    @property
    def clamp_mode(self) -> int:
        return CLAMP_MODES[self.clamp_type]

    def lightslice_params(self, width, height, n_emitted, seed=2013, tile=32, map_resolution=32, kd_pivots=0):
        """keyword arguments of renzoku_amd.LightSlice (rzk_lightslice_params): the slices are the
        reference's kd-tree slices of at most slice_size gather points when slice_size > 0;
        kd_pivots=1 draws the tree's pivots from the LightSlice's MT stream in the
        reference's recursion order (a sequential host build), 0 from per-node streams"""
        if self.integrator != "vpl":
            raise RenderConfigError("LightSlice parameters come from a vpl integrator")
        ls = self.lightslice
        return dict(width=width, height=height, tile=tile, num_clusters=ls.clusters,
                    num_neighbor_slices=ls.neighbor_slices, map_resolution=map_resolution,
                    clamp_mode=self.clamp_mode, clamp=self.clamp_threshold,
                    incident_clamp=self.incident_clamp_threshold, n_emitted=n_emitted, seed=seed,
                    max_slice_clusters=ls.max_clusters,
                    kd_max_slice_size=ls.slice_size, kd_pivots=int(kd_pivots))

    def path_params(self, spp=None, seed=2013, tmin=1e-3):
        """keyword arguments of LightSlice.render_guided / Guide.render (rzk_pt_params)"""
        if self.integrator == "path":
            return dict(spp=spp or self.samples, max_bounce=self.max_bounce, direct_lighting=1, guide_mode=0,
                        seed=seed, tmin=tmin)
        return dict(spp=spp or 1, max_bounce=self.max_bounce, direct_lighting=int(self.direct_lighting),
                    guide_mode=1, seed=seed, tmin=tmin)


def config_file(path: str) -> str:
    """RenderConfigImporter::load's file name rule"""
    if path.endswith(".json"):
        return path
    path = path.rstrip("/")
    name = os.path.basename(path)
    return os.path.join(path, name + ".json")


def _get(obj, key, kind, where):
    if key not in obj:
        raise RenderConfigError(f"{where}: missing \"{key}\"")
    v = obj[key]
    if kind is bool:
        if not isinstance(v, bool):
            raise RenderConfigError(f"{where}: \"{key}\" must be true or false")
        return v
    if kind is int:
        if isinstance(v, bool) or not isinstance(v, (int, float)) or int(v) != v:
            raise RenderConfigError(f"{where}: \"{key}\" must be an integer")
        return int(v)
    if kind is float:
        if isinstance(v, bool) or not isinstance(v, (int, float)):
            raise RenderConfigError(f"{where}: \"{key}\" must be a number")
        return float(v)
    if not isinstance(v, kind):
        raise RenderConfigError(f"{where}: \"{key}\" has the wrong type")
    return v


def parse_render_config(doc: dict) -> RenderConfig:
    """load_integrator (render_config_importer.cc:74-145) on a parsed JSON document"""
    if not isinstance(doc, dict) or "integrator" not in doc or not isinstance(doc["integrator"], dict):
        raise RenderConfigError("render config: no \"integrator\" object")
    spec = doc["integrator"]
    max_bounce = _get(spec, "max_bounce", int, "integrator")
    kind = _get(spec, "type", str, "integrator")
    if kind == "path":
        return RenderConfig("path", max_bounce, samples=_get(spec, "samples", int, "integrator"))
    if kind != "vpl":
        raise RenderConfigError(f"invalid integrator type: {kind}")
    cfg = RenderConfig("vpl", max_bounce, direct_lighting=_get(spec, "direct_lighting", bool, "integrator"))
    est = _get(spec, "radiance_estimate", str, "integrator")
    if est not in ESTIMATES:
        raise RenderConfigError(f"unsupported radiance_estimate: {est}")
    cfg.radiance_estimate = est
    ls = _get(spec, "lightslice", dict, "integrator")
    cfg.lightslice = LightSliceConfig(clusters=_get(ls, "clusters", int, "lightslice"),
                                      max_clusters=_get(ls, "max_clusters", int, "lightslice"),
                                      slice_size=_get(ls, "slice_size", int, "lightslice"),
                                      density_radius=_get(ls, "density_radius", float, "lightslice"),
                                      neighbor_slices=_get(ls, "neighbor_slices", int, "lightslice"))
    cfg.start_particles = _get(spec, "start_particles", int, "integrator")
    clamp = _get(spec, "clamp_type", str, "integrator")
    if clamp not in CLAMP_MODES:
        raise RenderConfigError(f"unsupported clamp_type: {clamp}")
    cfg.clamp_type = clamp
    thr = _get(spec, "clamp_threshold", dict, "integrator")
    if clamp == "radiance":
        cfg.clamp_threshold = _get(thr, "radiance", float, "clamp_threshold")
        cfg.incident_clamp_threshold = _get(thr, "incident_radiance", float, "clamp_threshold")
    elif clamp != "none":
        cfg.clamp_threshold = _get(thr, clamp, float, "clamp_threshold")
        cfg.incident_clamp_threshold = _get(thr, "incident_" + clamp, float, "clamp_threshold")
    cfg.clamp_relaxation = _get(spec, "clamp_relaxation", float, "integrator")
    return cfg


def load_render_config(path: str) -> RenderConfig:
    """RenderConfigImporter::load + load_integrator for a file or a configuration directory"""
    with open(config_file(path)) as f:
        try:
            doc = json.load(f)
        except json.JSONDecodeError as e:
            raise RenderConfigError(f"{config_file(path)}: {e}") from None
    return parse_render_config(doc)
