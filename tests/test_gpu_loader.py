"""A scene read from the XML/OBJ format (tests/golden/scenes/room.xml) goes through the
drop-in host: packed with the host MTBVH or with the MTBVH built on the device, it
renders the oracle's frame bit for bit with the naive and the global-queue tracers."""
import os

import numpy as np
import pytest

from zillumgl_amd import loader, render, scene, _lib
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

ROOM = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "scenes", "room.xml")


def test_xml_scene_renders_oracle_frame():
    desc = loader.load_scene(ROOM)
    ph = scene.pack(desc)
    pd = scene.pack(desc, bvh="device")
    c = render.Context(0)
    try:
        for integ in (_lib.INTEGRATOR_NAIVE, _lib.INTEGRATOR_GLOBAL_QUEUE):
            ref, _ = orc.render_pixels(ph, orc.params(integ, 0, 4, 1, 5))
            for name, ps in (("host", ph), ("device", pd)):
                print(f"[gpu] loader: room.xml, integrator {integ}, {name} MTBVH", flush=True)
                c.upload(ps)
                c.clear()
                c.render(integ, 0, 4, 1, 5)
                gpu = c.read_frame()[..., :3]
                assert np.array_equal(gpu.view(np.uint32), ref.view(np.uint32)), (integ, name)
            assert ref.mean() > 0.01
    finally:
        c.close()
