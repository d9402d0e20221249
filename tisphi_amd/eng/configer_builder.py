"""Scene-JSON loader with the interface of the reference's ``SimConfiger``.

Reference: eng/configer_builder.py:3-39.  The JSON layout (``Configuration``,
``Materials``, ``Blocks``, ``Bodies``, ``Motions``) is accepted unchanged;
section getters return an empty list when the section is absent, and
``get_cfg(name, enforce_exist=True)`` raises ``AssertionError`` for a missing
key exactly like the reference (a missing key otherwise raises ``KeyError``).
"""
import copy
import json


class SimConfiger:
    def __init__(self, scene_file_path=None, config=None) -> None:
        """Load a scene from ``scene_file_path`` or take an already-parsed
        ``config`` dict (used by batched / served scenes generated in code)."""
        self.scene_file_path = scene_file_path
        if config is not None:
            self.config = copy.deepcopy(config)
        else:
            with open(scene_file_path, "r") as f:
                self.config = json.load(f)

    def get_cfg(self, name, enforce_exist=False):
        if enforce_exist:
            assert name in self.config["Configuration"], f"missing key {name!r}"
        return self.config["Configuration"][name]

    def get_cfg_default(self, name, default):
        """Configuration value or ``default`` (for the optional engine keys)."""
        return self.config["Configuration"].get(name, default)

    def _section(self, key):
        return self.config[key] if key in self.config else []

    def get_materials(self):
        return self._section("Materials")

    def get_blocks(self):
        return self._section("Blocks")

    def get_bodies(self):
        return self._section("Bodies")

    def get_motions(self):
        return self._section("Motions")
