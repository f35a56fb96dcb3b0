"""Colour look-up tables for the display colour map.

Mirror of the reference's LutData / LutManager (src/lut_manager.rs:4-18,
150-193): a LUT is 768 bytes, red[256] then green[256] then blue[256];
`reverse()` reverses each channel.  The reference embeds its .lut files in
the binary; here a LutManager lists and loads the .lut files of a directory.
"""
from __future__ import annotations

import os
from dataclasses import dataclass
from typing import List


@dataclass
class LutData:
    name: str
    red: bytes
    green: bytes
    blue: bytes

    @classmethod
    def from_bytes(cls, name: str, data: bytes) -> "LutData":
        if len(data) != 768:  # src/lut_manager.rs:168-174
            raise ValueError("Invalid LUT file size")
        return cls(name, bytes(data[0:256]), bytes(data[256:512]), bytes(data[512:768]))

    def reverse(self) -> None:  # src/lut_manager.rs:13-17
        self.red, self.green, self.blue = self.red[::-1], self.green[::-1], self.blue[::-1]

    def to_bytes(self) -> bytes:
        """The 768-byte layout the renderer takes (src/main.rs:330-334)."""
        return self.red + self.green + self.blue


class LutManager:
    def __init__(self, directory: str):
        self.directory = directory

    def get_available_luts(self) -> List[str]:
        names = [f[:-4] for f in os.listdir(self.directory) if f.endswith(".lut")]
        return sorted(names)

    def load_lut(self, name: str) -> LutData:
        path = os.path.join(self.directory, name + ".lut")
        if not os.path.exists(path):
            raise FileNotFoundError(f"LUT '{name}' not found")
        with open(path, "rb") as f:
            return LutData.from_bytes(name, f.read())


def gray() -> LutData:
    """A linear gray ramp (entry i = (i, i, i))."""
    r = bytes(range(256))
    return LutData("gray", r, r, r)
