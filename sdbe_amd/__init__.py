"""MI355X (gfx950) SWARM data back-end GPU stage.

``sdbe_amd.stage.SdbeStage`` drives the C-ABI of include/sdbe_stage.h
(libsdbe.so); ``sdbe_amd.synth`` generates seeded B-engine VDIF input.
"""
__all__ = ["stage", "synth", "build"]
