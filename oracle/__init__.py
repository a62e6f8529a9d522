"""TEST INFRASTRUCTURE ONLY: CPU FP64 restatement of the reference path."""
