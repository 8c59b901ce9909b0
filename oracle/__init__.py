"""CPU oracle for the cudacompress_amd parity tests (test infrastructure, never imported by the product)."""
