"""The reference's demonstrations of the acoustic path, run on the device
through this package's drop-in API (simulation/examples/run_fwi_demo.py,
run_fwi_demo_3d.py, run_pipeline_demo.py, wavefield_snapshot.py):

    python -m openusct_amd.examples.fwi_2d      # ring FWI of a flaw, 2D
    python -m openusct_amd.examples.fwi_3d      # cylinder-array FWI of a spherical flaw
    python -m openusct_amd.examples.pipeline    # simulate -> HDF5 dataset -> plugin FWI
    python -m openusct_amd.examples.wavefield   # pressure snapshots of one transmit

Each ``main(out_dir=..., ...)`` returns its numbers (misfit history, the
recovered sound speed at the flaw) and saves the reference's figure when
matplotlib is importable.
"""
