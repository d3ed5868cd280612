"""ORACLE -- TEST INFRASTRUCTURE ONLY (CPU restatement of the reference path)."""
