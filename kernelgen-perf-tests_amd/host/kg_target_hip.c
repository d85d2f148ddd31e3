/* kg_target_hip.c -- marks the `hip` target: the per-test programs run the
 * device phases (init, alloc, load, save, free) through libkgstencil.so. */
const int kg_target_has_device = 1;
