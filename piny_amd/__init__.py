"""MI355X-native force evaluation for path-integral water, a drop-in for
PINY_MD's classical force call (energy_control / energy_control_pimd)."""
