/* Numeric tables of the spectral upsampling, embedded from the package's
 * data/ files so the shared library is self-contained:
 *   spectrum_mlp.bin  RGB/lambda MLP weights (16096 float32)
 *   cie_tables.bin    D65 SPD (421) + CIE 1931 2-degree CMF (421 x 3) */
    .section .rodata
    .balign 64
    .global ap_spectrum_mlp_bin
    .hidden ap_spectrum_mlp_bin
ap_spectrum_mlp_bin:
    .incbin "spectrum_mlp.bin"
    .global ap_spectrum_mlp_bin_end
    .hidden ap_spectrum_mlp_bin_end
ap_spectrum_mlp_bin_end:
    .balign 64
    .global ap_cie_tables_bin
    .hidden ap_cie_tables_bin
ap_cie_tables_bin:
    .incbin "cie_tables.bin"
    .global ap_cie_tables_bin_end
    .hidden ap_cie_tables_bin_end
ap_cie_tables_bin_end:
    .section .note.GNU-stack,"",@progbits
