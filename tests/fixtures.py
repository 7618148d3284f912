"""Loaders of the reference's bundled fixtures (tests/golden/voreen).

nucleon.{vvd,raw}, walnut.{vvd,raw}: voreen/resource/voreenve/volumes
nucleon.tfi, walnut.tfi:             voreen/resource/voreencore/transferfuncs
The .vvd format is VvdVolumeReader's XML (RawData filename/format/x/y/z,
Offset/Spacing meta items, VolumeHash/VolumeMinMax derived items); the .tfi
format is TransFuncIntensity XML (alphaMode, texDimensions, gammaValue,
domain, threshold, keys with colorL/colorR and split).
"""
import os
import xml.etree.ElementTree as ET

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "voreen")


def load_vvd(name):
    root = ET.parse(os.path.join(GOLDEN, name + ".vvd")).getroot()
    raw = root.find(".//RawData")
    dims = tuple(int(raw.get(k)) for k in ("x", "y", "z"))
    dtype = {"uint8": np.uint8, "uint16": np.uint16}[raw.get("format")]
    meta = {}
    for item in root.iter("MetaItem"):
        v = item.find("value")
        if v is not None and v.get("x") is not None:
            meta[item.get("name")] = tuple(float(v.get(k)) for k in ("x", "y", "z"))
    derived = {}
    for item in root.iter("DerivedItem"):
        if item.get("type") == "VolumeHash":
            derived["hash"] = item.get("hash")
        elif item.get("type") == "VolumeMinMax":
            derived["min"] = float(item.get("minNorm"))
            derived["max"] = float(item.get("maxNorm"))
    data = np.fromfile(os.path.join(GOLDEN, raw.get("filename")), dtype=dtype)
    data = data.reshape(dims[2], dims[1], dims[0])
    return dict(data=data, dims=dims, spacing=meta.get("Spacing", (1.0, 1.0, 1.0)),
                offset=meta.get("Offset", (0.0, 0.0, 0.0)), derived=derived,
                raw_path=os.path.join(GOLDEN, raw.get("filename")))


def load_tfi(name):
    root = ET.parse(os.path.join(GOLDEN, name + ".tfi")).getroot()
    tf = root.find("TransFuncIntensity")

    def attr(tag, k="value"):
        return tf.find(tag).get(k)

    keys = []
    for key in tf.find("Keys"):
        inten = float(key.find("intensity").get("value"))
        split = key.find("split").get("value") == "true"
        cl = key.find("colorL")
        col_l = tuple(int(cl.get(c)) for c in "rgba")
        cr = key.find("colorR")
        col_r = tuple(int(cr.get(c)) for c in "rgba") if (split and cr is not None) else col_l
        keys.append((inten, col_l, col_r) if split else (inten, col_l))
    return dict(keys=keys, width=int(attr("texDimensions", "x")),
                alpha_mode=int(attr("alphaMode")), gamma=float(attr("gammaValue")),
                domain=(float(attr("domain", "x")), float(attr("domain", "y"))),
                threshold=(float(attr("threshold", "x")), float(attr("threshold", "y"))))
