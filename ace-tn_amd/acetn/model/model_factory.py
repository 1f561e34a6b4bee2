"""Model registry (acetn/model/model_factory.py:4-83)."""
from .model import Model
from .models import models


class ModelFactory:
    def __init__(self):
        self._models = dict(models)

    def register(self, name, cls):
        if not issubclass(cls, Model):
            raise TypeError("Model class must inherit from Model base class")
        self._models[name] = cls

    def create(self, config):
        if config.name is None:
            raise ValueError("Model has not been set.")
        if config.name not in self._models:
            raise KeyError(f"Model name: '{config.name}' is not registered.")
        return self._models[config.name](config)

    def list_registered(self):
        return list(self._models.keys())


model_factory = ModelFactory()
