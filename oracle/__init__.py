"""CPU oracle for the fixed-trajectory NUFFT / CG-SENSE path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product (``trajsparsenufft_amd``)
may import, link or execute anything under ``oracle/``; only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` use
it, and then only as the checker or as the timed CPU baseline.

Modules
-------
``tkbn``     Kaiser-Bessel interpolation coefficients and apodization, a
             restatement of torchkbnufft 1.5.2 (third-party, not vendored in
             the reference; called at reference
             ``src/export_torchkbnufft_gpu_case.py:537-542``).
``nudft``    FP64 direct non-uniform DFT with SENSE, and an FP64 CG-SENSE on it.
``layout``   Pack layout: row grouping, column assignment, 2:4 arrangement,
             split24, compression/metadata, header, dense expansion.
``solver``   Component-wise arithmetic model (FP16 matrix x split2 panels),
             FP32 CSR operator and the reference CG update order.
``cpu_recon`` FP32 CSR + scipy.fft CG-SENSE, the timed CPU baseline.

Parity status of each function is stated in its docstring ("pinned by ...",
or "parity unpinned").
"""
