"""przm_amd: the PRZM GPU encode path for MI355X (gfx950).

gpu      -- the reference `gpu` package API (GPUCompressor, IsAvailable, ...)
serve    -- continuous-batching serving loop
datagen  -- seeded synthetic inputs
"""
__all__ = ["gpu", "serve", "datagen"]
