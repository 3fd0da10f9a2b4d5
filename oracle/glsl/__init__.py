"""llvmpipe harness for the reference GLSL (test infrastructure)."""
